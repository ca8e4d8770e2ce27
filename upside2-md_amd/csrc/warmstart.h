// Host side of the BP warm-start network: device copies of its parameters in
// the kernels' layouts (MlpParams), and the trainer (MlpTrainer: data on the
// device, full-data loss and gradient, Adam steps on minibatches).  Every
// product runs on the matrix cores through the kernels of mlp_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>
#include "dev_mlp.h"

hipError_t upside_mlp_launch_partial(const upside::MlpDev& w, const float* feat, const float* tgt, int n_rows,
                                     int row0, int n_batch, float inv_n, float* partial, int n_part,
                                     hipStream_t s);
hipError_t upside_mlp_launch_reduce(const float* partial, int n_part, float* master, float* m1, float* m2,
                                    float* grad_out, int adam, float lr, float beta1, float beta2, float eps,
                                    float bc1, float bc2, hipStream_t s);
hipError_t upside_mlp_launch_pack(const float* master, __bf16* w1, __bf16* w2p, __bf16* w3p, __bf16* w3tp,
                                  __bf16* w2tp, float* b1, float* b2, float* b3, hipStream_t s);
hipError_t upside_mlp_launch_forward(const upside::MlpDev& w, const float* feat, int n, float* out,
                                     hipStream_t s);
int upside_mlp_partial_size();

namespace upside {

// The network's parameters on one device: the float32 master vector and the
// bf16/float32 arrays the kernels read (one allocation, 256-byte aligned
// pieces).
struct MlpParams {
    char* base = nullptr;
    float* master = nullptr;
    MlpDev dev{};
    MlpParams();
    ~MlpParams();
    MlpParams(const MlpParams&) = delete;
    MlpParams& operator=(const MlpParams&) = delete;
    // copy a host parameter vector (MLP_NPARAM floats) and repack
    void set(const float* p, hipStream_t s);
    // repack after the master vector changed on the device
    void pack(hipStream_t s);
    void get(float* p) const;
};

}  // namespace upside
