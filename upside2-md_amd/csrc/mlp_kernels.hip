// Training and inference kernels of the BP warm-start network (dev_mlp.h):
// forward and backward passes on the matrix cores (bf16 operands, f32
// accumulation), weight gradients as matrix-core products over the rows,
// fixed-order reduction of the per-workgroup partial gradients, Adam, and the
// packing of the float32 master parameters into the kernels' bf16 layouts.
//
// Loss of one residue row (the network predicts corrections delta of the
// first BP pass's log beliefs): z_r = f[6 + r] + delta_r over its valid
// states (mask f[18 + r]), p = softmax(z), target t = the converged BP
// marginal; loss = -sum_r t_r log p_r, averaged over the rows of the batch.
// d loss / d delta_r = (p_r - t_r) / n_rows.
#include <hip/hip_runtime.h>
#include <cstdio>
#include "dev_mlp.h"

namespace upside {

constexpr int TR_WAVES = 4;                 // waves per training workgroup
constexpr int TR_THREADS = 64 * TR_WAVES;
constexpr int TR_ROWS = 16 * TR_WAVES;      // rows per chunk (one 16-row tile per wave)
constexpr int TR_KB = TR_ROWS / 32;         // 32-row K blocks per chunk
// weight-gradient output blocks (16 x 16): W1 8 x 2, W2 8 x 8, W3 1 x 8
constexpr int TR_BLK1 = (MLP_H / 16) * (MLP_IN / 16);
constexpr int TR_BLK2 = (MLP_H / 16) * (MLP_H / 16);
constexpr int TR_BLK3 = (MLP_OUT / 16) * (MLP_H / 16);
constexpr int TR_NBLK = TR_BLK1 + TR_BLK2 + TR_BLK3;
constexpr int TR_BLK_PER_WAVE = (TR_NBLK + TR_WAVES - 1) / TR_WAVES;
constexpr int TR_PARTIAL = MLP_NPARAM + 1;  // gradients, then the loss

// Per-row loss and its gradient for the row held by lanes c (outputs 0..3)
// and c + 16 (outputs 4..5): y in D layout; writes the gradient into dy (D
// layout, zero beyond the valid states) and returns the row's loss on lane c
// (0 elsewhere).
__device__ __forceinline__ float row_loss_grad(const float* feat, const float* tgt, int n, int row, f32x4 y,
                                               float inv_n, f32x4& dy) {
    const int l = (int)(threadIdx.x & 63), g = l >> 4, c = l & 15;
    // gather the row's six outputs to every lane of its two groups
    float yo[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        yo[i] = __shfl(y[i], c, 64);
        yo[4 + i] = __shfl(y[i], 16 + c, 64);
    }
    dy = f32x4{0.f, 0.f, 0.f, 0.f};
    if (row >= n || g > 1) return 0.f;
    float z[MLP_USED], p[MLP_USED], mx = -3.0e38f;
    bool ok[MLP_USED];
#pragma unroll
    for (int r = 0; r < MLP_USED; ++r) {
        ok[r] = feat[row * MLP_IN + 18 + r] > 0.5f;
        z[r] = feat[row * MLP_IN + 6 + r] + yo[r];
        if (ok[r]) mx = fmaxf(mx, z[r]);
    }
    float s = 0.f;
#pragma unroll
    for (int r = 0; r < MLP_USED; ++r) {
        p[r] = ok[r] ? expf(z[r] - mx) : 0.f;
        s += p[r];
    }
    const float lse = mx + logf(s), is = 1.f / s;
    float loss = 0.f;
#pragma unroll
    for (int r = 0; r < MLP_USED; ++r) {
        const float t = tgt[row * 8 + r];
        if (ok[r]) loss -= t * (z[r] - lse);
        p[r] = ok[r] ? (p[r] * is - t) * inv_n : 0.f;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int o = 4 * g + i;
        dy[i] = o < MLP_USED ? p[o] : 0.f;
    }
    return g == 0 ? loss * inv_n : 0.f;
}

// Store a D-layout block (outputs blk*16 + 4g + i of tile row c) into a
// transposed LDS array T[feature][TR_ROWS] as bf16.
__device__ __forceinline__ void store_t(__bf16* T, int blk, f32x4 v, int wave) {
    const int l = (int)(threadIdx.x & 63), g = l >> 4, c = l & 15;
#pragma unroll
    for (int i = 0; i < 4; ++i) T[(blk * 16 + 4 * g + i) * TR_ROWS + wave * 16 + c] = (__bf16)v[i];
}

// Partial gradients of the loss over the rows [row0, row0 + n_batch) (taken
// modulo n_rows) handled by this workgroup: chunks b, b + gridDim.x, ... of
// TR_ROWS rows.  partial[blockIdx.x][TR_PARTIAL].
__global__ void __launch_bounds__(TR_THREADS) mlp_train_partial_kernel(MlpDev w, const float* feat,
                                                                        const float* tgt, int n_rows, int row0,
                                                                        int n_batch, float inv_n, float* partial) {
    __shared__ __attribute__((aligned(16))) __bf16 XT[MLP_IN * TR_ROWS];
    __shared__ __attribute__((aligned(16))) __bf16 H1T[MLP_H * TR_ROWS];
    __shared__ __attribute__((aligned(16))) __bf16 H2T[MLP_H * TR_ROWS];
    __shared__ __attribute__((aligned(16))) __bf16 DYT[MLP_OUT * TR_ROWS];
    __shared__ __attribute__((aligned(16))) __bf16 DH2T[MLP_H * TR_ROWS];
    __shared__ __attribute__((aligned(16))) __bf16 DH1T[MLP_H * TR_ROWS];
    __shared__ float lred[TR_WAVES];
    const int wave = (int)threadIdx.x >> 6, l = (int)threadIdx.x & 63, g = l >> 4, c = l & 15;
    const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
    f32x4 acc[TR_BLK_PER_WAVE];
#pragma unroll
    for (int k = 0; k < TR_BLK_PER_WAVE; ++k) acc[k] = z4;
    float bacc = 0.f, bacc2 = 0.f;  // bias gradients: b1/b2 entry of this thread, b3 entry (tid < 16)
    float loss = 0.f;
    const int n_chunks = (n_batch + TR_ROWS - 1) / TR_ROWS;
    for (int ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {
        // one 16-row tile per wave: rows of the batch, wrapped into the data
        const int tb = ch * TR_ROWS + wave * 16;   // batch position of the tile's first row
        const int bpos = tb + c;
        const bool live = bpos < n_batch;
        const int row = live ? (row0 + bpos) % n_rows : 0;
        // forward (as mlp_tile_forward, on this lane's wrapped row), keeping
        // the inputs transposed in LDS for the layer-1 weight gradient
        f32x4 h1[MLP_H / 16], h2[MLP_H / 16], y;
        {
            bf16x8 x;
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = (__bf16)(live ? feat[row * MLP_IN + 8 * g + j] : 0.f);
#pragma unroll
            for (int j = 0; j < 8; ++j) XT[(8 * g + j) * TR_ROWS + wave * 16 + c] = x[j];
#pragma unroll
            for (int ob = 0; ob < MLP_H / 16; ++ob) {
                f32x4 a = mfma_bf16(ld_bf16x8(w.w1 + (ob * 16 + c) * MLP_IN + 8 * g), x, z4);
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = fmaxf(a[i] + w.b1[ob * 16 + 4 * g + i], 0.f);
                h1[ob] = a;
            }
            bf16x8 b[MLP_H / 32];
#pragma unroll
            for (int kb = 0; kb < MLP_H / 32; ++kb) b[kb] = pack_blocks(h1[2 * kb], h1[2 * kb + 1]);
#pragma unroll
            for (int ob = 0; ob < MLP_H / 16; ++ob) {
                f32x4 a = z4;
#pragma unroll
                for (int kb = 0; kb < MLP_H / 32; ++kb)
                    a = mfma_bf16(ld_bf16x8(w.w2p + (ob * 16 + c) * MLP_H + kb * 32 + 8 * g), b[kb], a);
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = fmaxf(a[i] + w.b2[ob * 16 + 4 * g + i], 0.f);
                h2[ob] = a;
            }
#pragma unroll
            for (int kb = 0; kb < MLP_H / 32; ++kb) b[kb] = pack_blocks(h2[2 * kb], h2[2 * kb + 1]);
            f32x4 a = z4;
#pragma unroll
            for (int kb = 0; kb < MLP_H / 32; ++kb) a = mfma_bf16(ld_bf16x8(w.w3p + c * MLP_H + kb * 32 + 8 * g), b[kb], a);
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] += w.b3[4 * g + i];
            y = a;
        }
        if (!live) {  // rows past the batch contribute nothing
#pragma unroll
            for (int ob = 0; ob < MLP_H / 16; ++ob) { h1[ob] = z4; h2[ob] = z4; }
        }
        // loss and output gradient (a dead row reads row 0 and is zeroed)
        f32x4 dy;
        {
            float lz = row_loss_grad(feat, tgt, n_rows, row, y, inv_n, dy);
            if (!live) { dy = z4; lz = 0.f; }
            loss += lz;
        }
        // backward through layer 3: dH2^T = W3^T dY^T, masked by ReLU
        f32x4 dh2[MLP_H / 16], dh1[MLP_H / 16];
        {
            const bf16x8 b = pack_blocks(dy, z4);
#pragma unroll
            for (int jb = 0; jb < MLP_H / 16; ++jb) {
                f32x4 a = mfma_bf16(ld_bf16x8(w.w3tp + (jb * 16 + c) * 32 + 8 * g), b, z4);
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = h2[jb][i] > 0.f ? a[i] : 0.f;
                dh2[jb] = a;
            }
            bf16x8 bb[MLP_H / 32];
#pragma unroll
            for (int kb = 0; kb < MLP_H / 32; ++kb) bb[kb] = pack_blocks(dh2[2 * kb], dh2[2 * kb + 1]);
#pragma unroll
            for (int ib = 0; ib < MLP_H / 16; ++ib) {
                f32x4 a = z4;
#pragma unroll
                for (int kb = 0; kb < MLP_H / 32; ++kb)
                    a = mfma_bf16(ld_bf16x8(w.w2tp + (ib * 16 + c) * MLP_H + kb * 32 + 8 * g), bb[kb], a);
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = h1[ib][i] > 0.f ? a[i] : 0.f;
                dh1[ib] = a;
            }
        }
        // transposed bf16 copies for the weight gradients
#pragma unroll
        for (int ob = 0; ob < MLP_H / 16; ++ob) {
            store_t(H1T, ob, h1[ob], wave);
            store_t(H2T, ob, h2[ob], wave);
            store_t(DH2T, ob, dh2[ob], wave);
            store_t(DH1T, ob, dh1[ob], wave);
        }
        store_t(DYT, 0, dy, wave);
        __syncthreads();
        // weight gradients: dW[o][i] = sum over the chunk's rows of dH[row][o] * In[row][i]
#pragma unroll
        for (int k = 0; k < TR_BLK_PER_WAVE; ++k) {
            const int blk = wave + k * TR_WAVES;
            if (blk >= TR_NBLK) break;
            const __bf16 *A, *B;
            int mb, nb;
            if (blk < TR_BLK1) { mb = blk / (MLP_IN / 16); nb = blk % (MLP_IN / 16); A = DH1T; B = XT; }
            else if (blk < TR_BLK1 + TR_BLK2) { const int q = blk - TR_BLK1; mb = q / (MLP_H / 16); nb = q % (MLP_H / 16); A = DH2T; B = H1T; }
            else { const int q = blk - TR_BLK1 - TR_BLK2; mb = 0; nb = q; A = DYT; B = H2T; }
#pragma unroll
            for (int kb = 0; kb < TR_KB; ++kb)
                acc[k] = mfma_bf16(ld_bf16x8(A + (mb * 16 + c) * TR_ROWS + kb * 32 + 8 * g),
                                   ld_bf16x8(B + (nb * 16 + c) * TR_ROWS + kb * 32 + 8 * g), acc[k]);
        }
        // bias gradients: b1 (128), b2 (128), b3 (16) from the same bf16 values, rows in order
        {
            const int t = (int)threadIdx.x;  // 0..255: b1 [0,128), b2 [128,256); t < 16 also b3
            const __bf16* src = t < MLP_H ? DH1T + t * TR_ROWS : DH2T + (t - MLP_H) * TR_ROWS;
            for (int r = 0; r < TR_ROWS; ++r) bacc += (float)src[r];
            if (t < MLP_OUT)
                for (int r = 0; r < TR_ROWS; ++r) bacc2 += (float)DYT[t * TR_ROWS + r];
        }
        __syncthreads();
    }
    // write the partial gradients of this workgroup
    float* out = partial + (size_t)blockIdx.x * TR_PARTIAL;
#pragma unroll
    for (int k = 0; k < TR_BLK_PER_WAVE; ++k) {
        const int blk = wave + k * TR_WAVES;
        if (blk >= TR_NBLK) break;
        int base, ld, mb, nb;
        if (blk < TR_BLK1) { mb = blk / (MLP_IN / 16); nb = blk % (MLP_IN / 16); base = MLP_OFF_W1; ld = MLP_IN; }
        else if (blk < TR_BLK1 + TR_BLK2) { const int q = blk - TR_BLK1; mb = q / (MLP_H / 16); nb = q % (MLP_H / 16); base = MLP_OFF_W2; ld = MLP_H; }
        else { const int q = blk - TR_BLK1 - TR_BLK2; mb = 0; nb = q; base = MLP_OFF_W3; ld = MLP_H; }
#pragma unroll
        for (int i = 0; i < 4; ++i) out[base + (mb * 16 + 4 * g + i) * ld + nb * 16 + c] = acc[k][i];
    }
    {
        const int t = (int)threadIdx.x;
        if (t < MLP_H) out[MLP_OFF_B1 + t] = bacc;
        else out[MLP_OFF_B2 + t - MLP_H] = bacc;
        if (t < MLP_OUT) out[MLP_OFF_B3 + t] = bacc2;
    }
    // the loss: lanes in order, then waves in order
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    if (l == 0) lred[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int k = 0; k < TR_WAVES; ++k) s += lred[k];
        out[MLP_NPARAM] = s;
    }
}

// Sum of the partial gradients over workgroups in workgroup order, then
// (adam != 0) one Adam step on the master parameters; grad_out (may be null)
// receives the summed gradient and loss.
__global__ void mlp_reduce_adam_kernel(const float* partial, int n_part, float* master, float* m1, float* m2,
                                       float* grad_out, int adam, float lr, float beta1, float beta2, float eps,
                                       float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > MLP_NPARAM) return;
    float gsum = 0.f;
    for (int b = 0; b < n_part; ++b) gsum += partial[(size_t)b * TR_PARTIAL + p];
    if (grad_out) grad_out[p] = gsum;
    if (!adam || p == MLP_NPARAM) return;
    const float a = beta1 * m1[p] + (1.f - beta1) * gsum;
    const float v = beta2 * m2[p] + (1.f - beta2) * gsum * gsum;
    m1[p] = a;
    m2[p] = v;
    master[p] -= lr * (a / bc1) / (sqrtf(v / bc2) + eps);
}

// bf16 layouts of the master parameters (one thread per destination entry).
__global__ void mlp_pack_kernel(const float* master, __bf16* w1, __bf16* w2p, __bf16* w3p, __bf16* w3tp,
                                __bf16* w2tp, float* b1, float* b2, float* b3) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const float* W1 = master + MLP_OFF_W1;
    const float* W2 = master + MLP_OFF_W2;
    const float* W3 = master + MLP_OFF_W3;
    if (t < MLP_H * MLP_IN) w1[t] = (__bf16)W1[t];
    if (t < MLP_H * MLP_H) {
        const int o = t / MLP_H, s = t % MLP_H, kb = s / 32, g = (s % 32) / 8, j = s % 8;
        w2p[t] = (__bf16)W2[o * MLP_H + mlp_perm(kb, g, j)];
        // w2tp row i, slot (kb, g, j) = W2[perm][i]
        w2tp[t] = (__bf16)W2[mlp_perm(kb, g, j) * MLP_H + o];
    }
    if (t < MLP_OUT * MLP_H) {
        const int o = t / MLP_H, s = t % MLP_H, kb = s / 32, g = (s % 32) / 8, j = s % 8;
        w3p[t] = (__bf16)W3[o * MLP_H + mlp_perm(kb, g, j)];
    }
    if (t < MLP_H * 32) {
        const int j = t / 32, s = t % 32, g = s / 8, jj = s % 8;
        w3tp[t] = jj < 4 ? (__bf16)W3[(4 * g + jj) * MLP_H + j] : (__bf16)0.f;
    }
    if (t < MLP_H) { b1[t] = master[MLP_OFF_B1 + t]; b2[t] = master[MLP_OFF_B2 + t]; }
    if (t < MLP_OUT) b3[t] = master[MLP_OFF_B3 + t];
}

// Outputs of every row: out[n][MLP_OUT]; one wave per 16-row tile.
__global__ void __launch_bounds__(256) mlp_forward_kernel(MlpDev w, const float* feat, int n, float* out) {
    const int wave = (int)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int r0 = wave * 16;
    if (r0 >= n) return;  // uniform over the wave
    const int l = (int)threadIdx.x & 63, g = l >> 4, c = l & 15;
    f32x4 h1[MLP_H / 16], h2[MLP_H / 16], y;
    mlp_tile_forward(w, feat, n, r0, h1, h2, y);
    if (r0 + c < n)
#pragma unroll
        for (int i = 0; i < 4; ++i) out[(r0 + c) * MLP_OUT + 4 * g + i] = y[i];
}

}  // namespace upside

// ---- host launchers (used by the trainer in warmstart_host.cpp) ------------------
using namespace upside;

hipError_t upside_mlp_launch_partial(const MlpDev& w, const float* feat, const float* tgt, int n_rows, int row0,
                                     int n_batch, float inv_n, float* partial, int n_part, hipStream_t s) {
    hipLaunchKernelGGL(mlp_train_partial_kernel, dim3(n_part), dim3(TR_THREADS), 0, s, w, feat, tgt, n_rows, row0,
                       n_batch, inv_n, partial);
    return hipGetLastError();
}

hipError_t upside_mlp_launch_reduce(const float* partial, int n_part, float* master, float* m1, float* m2,
                                    float* grad_out, int adam, float lr, float beta1, float beta2, float eps,
                                    float bc1, float bc2, hipStream_t s) {
    const int n = MLP_NPARAM + 1;
    hipLaunchKernelGGL(mlp_reduce_adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, partial, n_part, master, m1,
                       m2, grad_out, adam, lr, beta1, beta2, eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t upside_mlp_launch_pack(const float* master, __bf16* w1, __bf16* w2p, __bf16* w3p, __bf16* w3tp,
                                  __bf16* w2tp, float* b1, float* b2, float* b3, hipStream_t s) {
    const int n = MLP_H * MLP_H;
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((n + 255) / 256), dim3(256), 0, s, master, w1, w2p, w3p, w3tp, w2tp,
                       b1, b2, b3);
    return hipGetLastError();
}

hipError_t upside_mlp_launch_forward(const MlpDev& w, const float* feat, int n, float* out, hipStream_t s) {
    const int waves = (n + 15) / 16;
    const int blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(256), 0, s, w, feat, n, out);
    return hipGetLastError();
}

int upside_mlp_partial_size() { return TR_PARTIAL; }

// ---- loss and gradient without allocation ---------------------------------------
// Callers that cannot allocate (code under graph capture, serving loops) get
// the network's loss and gradient on device buffers they own, with every
// intermediate in a file-scope device array: the packed bf16 weights and the
// per-workgroup partial gradients.  Each GPU holds its own copy of the array
// from the library's first use on it until the process ends; its device
// address is looked up once per device.  Calls that overlap on one GPU share
// the array and race: callers serialise them.
namespace {
constexpr size_t NOALLOC_WEIGHTS = 96 * 1024;                               // bf16 copies and biases
constexpr size_t NOALLOC_BYTES = NOALLOC_WEIGHTS + (size_t)MLP_PARTS * TR_PARTIAL * sizeof(float);
constexpr int NOALLOC_MAX_ROWS = 1 << 24;                                    // row * MLP_IN stays in int range
}  // namespace

static __device__ __align__(256) char scratch[NOALLOC_BYTES];

namespace {
char* scratch_ptr[64];
bool scratch_ready[64];

// Device address of `scratch` on the current device (hipGetSymbolAddress once per device).
hipError_t scratch_address(char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    if (!scratch_ready[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_ptr[dev] = (char*)p;
        scratch_ready[dev] = true;
    }
    *out = scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

// Loss and gradient of the mean loss over n_rows rows, as upside_mlp_loss_grad
// computes them (same kernels, grid and order, so the same bits):
// params[MLP_NPARAM], features[n_rows][MLP_IN] and targets[n_rows][8] in
// device memory; out[MLP_NPARAM + 1] (gradient, then the loss) in device
// memory.  Runs on the null stream and returns when the result is written.
extern "C" int upside_mlp_loss_grad_noalloc(const float* params, const float* features, const float* targets,
                                            int n_rows, float* out) {
    if (n_rows <= 0 || n_rows > NOALLOC_MAX_ROWS) {
        std::fprintf(stderr, "ERROR: upside_mlp_loss_grad_noalloc: n_rows must be in [1, %d], got %d\n",
                     NOALLOC_MAX_ROWS, n_rows);
        return 1;
    }
    if (!params || !features || !targets || !out) {
        std::fprintf(stderr, "ERROR: upside_mlp_loss_grad_noalloc: null buffer\n");
        return 1;
    }
    char* base = nullptr;
    hipError_t e = scratch_address(&base);
    if (e != hipSuccess) {
        std::fprintf(stderr, "ERROR: upside_mlp_loss_grad_noalloc: scratch address: %s\n", hipGetErrorString(e));
        return 1;
    }
    // carve the weight copies (256-byte aligned pieces) and the partials
    char* p = base;
    auto take = [&](size_t b) { char* q = p; p += (b + 255) & ~(size_t)255; return q; };
    __bf16* w1 = (__bf16*)take(MLP_H * MLP_IN * 2);
    __bf16* w2p = (__bf16*)take(MLP_H * MLP_H * 2);
    __bf16* w3p = (__bf16*)take(MLP_OUT * MLP_H * 2);
    __bf16* w3tp = (__bf16*)take(MLP_H * 32 * 2);
    __bf16* w2tp = (__bf16*)take(MLP_H * MLP_H * 2);
    float* b1 = (float*)take(MLP_H * 4);
    float* b2 = (float*)take(MLP_H * 4);
    float* b3 = (float*)take(MLP_OUT * 4);
    float* partial = (float*)(base + NOALLOC_WEIGHTS);
    static_assert(((MLP_H * MLP_IN * 2 + 255) / 256 + (MLP_H * MLP_H * 2 + 255) / 256 * 2 +
                   (MLP_OUT * MLP_H * 2 + 255) / 256 + (MLP_H * 32 * 2 + 255) / 256 +
                   (MLP_H * 4 + 255) / 256 * 2 + (MLP_OUT * 4 + 255) / 256) * 256 <= NOALLOC_WEIGHTS,
                  "weight copies exceed their part of the scratch array");
    MlpDev w{w1, w2p, w3p, w3tp, w2tp, b1, b2, b3};
    const hipStream_t s = 0;
    if ((e = upside_mlp_launch_pack(params, w1, w2p, w3p, w3tp, w2tp, b1, b2, b3, s)) != hipSuccess ||
        (e = upside_mlp_launch_partial(w, features, targets, n_rows, 0, n_rows, 1.f / (float)n_rows, partial,
                                       MLP_PARTS, s)) != hipSuccess ||
        (e = upside_mlp_launch_reduce(partial, MLP_PARTS, nullptr, nullptr, nullptr, out, 0, 0.f, 0.f, 0.f, 0.f,
                                      1.f, 1.f, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess) {
        std::fprintf(stderr, "ERROR: upside_mlp_loss_grad_noalloc: %s\n", hipGetErrorString(e));
        return 1;
    }
    return 0;
}
