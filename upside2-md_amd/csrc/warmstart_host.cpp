// Trainer of the BP warm-start network and its C ABI (upside_mlp_*).  No
// counterpart in the reference.  The network, its loss and its layouts are
// described in dev_mlp.h and mlp_kernels.hip.
#include "warmstart.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/upside_engine.h"

using namespace upside;

namespace {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// splitmix64: the seeded generator of the initial weights
uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double unit(uint64_t& s) { return ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
double normal(uint64_t& s) { return std::sqrt(-2.0 * std::log(unit(s))) * std::cos(6.283185307179586 * unit(s)); }

}  // namespace

namespace upside {

MlpParams::MlpParams() {
    const size_t sizes[] = {MLP_NPARAM * 4, MLP_H * MLP_IN * 2, MLP_H * MLP_H * 2, MLP_OUT * MLP_H * 2,
                            MLP_H * 32 * 2, MLP_H * MLP_H * 2, MLP_H * 4, MLP_H * 4, MLP_OUT * 4};
    size_t total = 0;
    for (size_t b : sizes) total += up256(b);
    check(hipMalloc((void**)&base, total), "warm-start parameters");
    check(hipMemset(base, 0, total), "warm-start parameters");
    char* p = base;
    auto take = [&](size_t b) { char* q = p; p += up256(b); return q; };
    master = (float*)take(sizes[0]);
    dev.w1 = (const __bf16*)take(sizes[1]);
    dev.w2p = (const __bf16*)take(sizes[2]);
    dev.w3p = (const __bf16*)take(sizes[3]);
    dev.w3tp = (const __bf16*)take(sizes[4]);
    dev.w2tp = (const __bf16*)take(sizes[5]);
    dev.b1 = (const float*)take(sizes[6]);
    dev.b2 = (const float*)take(sizes[7]);
    dev.b3 = (const float*)take(sizes[8]);
}

MlpParams::~MlpParams() { (void)hipFree(base); }

void MlpParams::pack(hipStream_t s) {
    check(upside_mlp_launch_pack(master, (__bf16*)dev.w1, (__bf16*)dev.w2p, (__bf16*)dev.w3p, (__bf16*)dev.w3tp,
                                 (__bf16*)dev.w2tp, (float*)dev.b1, (float*)dev.b2, (float*)dev.b3, s),
          "warm-start pack");
}

void MlpParams::set(const float* p, hipStream_t s) {
    for (int k = 0; k < MLP_NPARAM; ++k)
        if (!std::isfinite(p[k])) throw std::runtime_error("warm-start parameters must be finite");
    check(hipMemcpyAsync(master, p, MLP_NPARAM * 4, hipMemcpyHostToDevice, s), "warm-start parameters");
    pack(s);
    check(hipStreamSynchronize(s), "warm-start parameters");
}

void MlpParams::get(float* p) const {
    check(hipMemcpy(p, master, MLP_NPARAM * 4, hipMemcpyDeviceToHost), "warm-start parameters");
}

}  // namespace upside

// The trainer: parameters, Adam moments, the training data and the partial
// gradient buffer, all on the device.
struct UpsideMlp {
    MlpParams params;
    float* m1 = nullptr;
    float* m2 = nullptr;
    float* grad = nullptr;     // MLP_NPARAM + 1 (loss last)
    float* feat = nullptr;     // [n_rows][MLP_IN]
    float* tgt = nullptr;      // [n_rows][8]
    float* partial = nullptr;  // [n_part][MLP_NPARAM + 1]
    int n_rows = 0;
    int n_part = MLP_PARTS;    // workgroups of the gradient kernel
    long long step = 0;
    hipStream_t stream = nullptr;

    explicit UpsideMlp(uint32_t seed) {
        check(hipStreamCreate(&stream), "stream");
        check(hipMalloc((void**)&m1, MLP_NPARAM * 4), "adam");
        check(hipMalloc((void**)&m2, MLP_NPARAM * 4), "adam");
        check(hipMemset(m1, 0, MLP_NPARAM * 4), "adam");
        check(hipMemset(m2, 0, MLP_NPARAM * 4), "adam");
        check(hipMalloc((void**)&grad, (MLP_NPARAM + 1) * 4), "gradient");
        check(hipMalloc((void**)&partial, (size_t)n_part * upside_mlp_partial_size() * 4), "partials");
        // He-normal hidden layers, zero output layer (the untrained network
        // leaves the first-pass beliefs unchanged), zero biases
        std::vector<float> p(MLP_NPARAM, 0.f);
        uint64_t s = 0x5EEDull ^ ((uint64_t)seed << 32);
        for (int k = 0; k < MLP_H * MLP_IN; ++k) p[MLP_OFF_W1 + k] = (float)(normal(s) * std::sqrt(2.0 / MLP_IN));
        for (int k = 0; k < MLP_H * MLP_H; ++k) p[MLP_OFF_W2 + k] = (float)(normal(s) * std::sqrt(2.0 / MLP_H));
        params.set(p.data(), stream);
    }
    ~UpsideMlp() {
        (void)hipFree(m1); (void)hipFree(m2); (void)hipFree(grad); (void)hipFree(partial);
        (void)hipFree(feat); (void)hipFree(tgt);
        (void)hipStreamDestroy(stream);
    }

    void set_data(const float* f, const float* t, int n) {
        if (n <= 0) throw std::runtime_error("no training rows");
        for (size_t k = 0; k < (size_t)n * MLP_IN; ++k)
            if (!std::isfinite(f[k])) throw std::runtime_error("training features must be finite");
        (void)hipFree(feat); (void)hipFree(tgt);
        feat = tgt = nullptr;
        check(hipMalloc((void**)&feat, (size_t)n * MLP_IN * 4), "training data");
        check(hipMalloc((void**)&tgt, (size_t)n * 8 * 4), "training data");
        check(hipMemcpy(feat, f, (size_t)n * MLP_IN * 4, hipMemcpyHostToDevice), "training data");
        check(hipMemcpy(tgt, t, (size_t)n * 8 * 4, hipMemcpyHostToDevice), "training data");
        n_rows = n;
    }

    // gradient of the mean loss over rows [row0, row0 + nb) (mod n_rows);
    // with adam, one Adam step
    void gradient(int row0, int nb, bool adam, float lr) {
        if (!n_rows) throw std::runtime_error("no training data set");
        check(upside_mlp_launch_partial(params.dev, feat, tgt, n_rows, row0, nb, 1.f / (float)nb, partial, n_part,
                                        stream), "warm-start gradient");
        float bc1 = 1.f, bc2 = 1.f;
        if (adam) {
            ++step;
            bc1 = 1.f - (float)std::pow(0.9, (double)step);
            bc2 = 1.f - (float)std::pow(0.999, (double)step);
        }
        check(upside_mlp_launch_reduce(partial, n_part, params.master, m1, m2, grad, adam ? 1 : 0, lr, 0.9f, 0.999f,
                                       1e-8f, bc1, bc2, stream), "warm-start reduce");
        if (adam) params.pack(stream);
    }
};

#define MLP_TRY try {
#define MLP_CATCH(rv)                                                   \
    } catch (const std::exception& e) {                                 \
        std::fprintf(stderr, "ERROR: %s\n", e.what());                  \
        return rv;                                                      \
    }

extern "C" {

UpsideMlp* upside_mlp_create(uint32_t seed) {
    MLP_TRY
    return new UpsideMlp(seed);
    MLP_CATCH(nullptr)
}

void upside_mlp_free(UpsideMlp* m) { delete m; }

int upside_mlp_n_param() { return MLP_NPARAM; }

int upside_mlp_get_params(UpsideMlp* m, float* p, int n) {
    MLP_TRY
    if (n != MLP_NPARAM) throw std::runtime_error("parameter vector has " + std::to_string(MLP_NPARAM) + " floats");
    m->params.get(p);
    return 0;
    MLP_CATCH(1)
}

int upside_mlp_set_params(UpsideMlp* m, const float* p, int n) {
    MLP_TRY
    if (n != MLP_NPARAM) throw std::runtime_error("parameter vector has " + std::to_string(MLP_NPARAM) + " floats");
    m->params.set(p, m->stream);
    return 0;
    MLP_CATCH(1)
}

int upside_mlp_set_data(UpsideMlp* m, const float* feat, const float* target, int n_rows) {
    MLP_TRY
    m->set_data(feat, target, n_rows);
    return 0;
    MLP_CATCH(1)
}

int upside_mlp_loss_grad(UpsideMlp* m, float* loss, float* grad) {
    MLP_TRY
    m->gradient(0, m->n_rows, false, 0.f);
    std::vector<float> g(MLP_NPARAM + 1);
    check(hipMemcpyAsync(g.data(), m->grad, g.size() * 4, hipMemcpyDeviceToHost, m->stream), "gradient");
    check(hipStreamSynchronize(m->stream), "gradient");
    if (loss) *loss = g[MLP_NPARAM];
    if (grad) std::memcpy(grad, g.data(), MLP_NPARAM * 4);
    return 0;
    MLP_CATCH(1)
}

int upside_mlp_train(UpsideMlp* m, int n_steps, int batch_rows, float lr, float* losses) {
    MLP_TRY
    if (n_steps < 0 || batch_rows <= 0 || !(lr > 0.f)) throw std::runtime_error("bad training arguments");
    if (!m->n_rows) throw std::runtime_error("no training data set");
    const int nb = batch_rows < m->n_rows ? batch_rows : m->n_rows;
    long long row0 = 0;
    for (int k = 0; k < n_steps; ++k) {
        m->gradient((int)(row0 % m->n_rows), nb, true, lr);
        row0 += nb;
        if (losses) {
            float l;
            check(hipMemcpyAsync(&l, m->grad + MLP_NPARAM, 4, hipMemcpyDeviceToHost, m->stream), "loss");
            check(hipStreamSynchronize(m->stream), "loss");
            losses[k] = l;
        }
    }
    check(hipStreamSynchronize(m->stream), "training");
    return 0;
    MLP_CATCH(1)
}

int upside_mlp_forward(UpsideMlp* m, const float* feat, int n_rows, float* out) {
    MLP_TRY
    if (n_rows <= 0) throw std::runtime_error("no rows");
    float *df = nullptr, *dout = nullptr;
    check(hipMalloc((void**)&df, (size_t)n_rows * MLP_IN * 4), "forward");
    check(hipMalloc((void**)&dout, (size_t)n_rows * MLP_OUT * 4), "forward");
    check(hipMemcpyAsync(df, feat, (size_t)n_rows * MLP_IN * 4, hipMemcpyHostToDevice, m->stream), "forward");
    check(upside_mlp_launch_forward(m->params.dev, df, n_rows, dout, m->stream), "forward");
    check(hipMemcpyAsync(out, dout, (size_t)n_rows * MLP_OUT * 4, hipMemcpyDeviceToHost, m->stream), "forward");
    check(hipStreamSynchronize(m->stream), "forward");
    (void)hipFree(df);
    (void)hipFree(dout);
    return 0;
    MLP_CATCH(1)
}

}  // extern "C"
