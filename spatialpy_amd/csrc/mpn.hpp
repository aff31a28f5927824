// Host side of the message-passing surrogate: parameters, activations,
// forward / backward / Adam on the GPU (kernels in mpn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "engine.hpp"

namespace ssa {

struct MpnSample {
    int n = 0, W = 0;
    float inv_h = 0.f;
    float *x = nullptr;      // [n][3]
    float *X0 = nullptr;     // [n][32]
    int *deg = nullptr, *nidx = nullptr;
    float *y = nullptr;      // [n][3] normalised targets
};

class Mpn {
  public:
    static constexpr int HID = 64, LAYERS = 3, NIN = 32, NOUT = 3;
    // flat FP32 parameter offsets
    static constexpr int OFF_WE = 0, OFF_BE = OFF_WE + NIN * HID;
    static constexpr int L0 = OFF_BE + HID;
    static constexpr int W1AB = 0, W1C = W1AB + HID * 128, B1 = W1C + 4 * HID, W2 = B1 + HID, B2 = W2 + HID * HID,
                         WN1 = B2 + HID, BN1 = WN1 + 128 * HID, WN2 = BN1 + HID, BN2 = WN2 + HID * HID,
                         LSIZE = BN2 + HID;
    static constexpr int OFF_WD = L0 + LAYERS * LSIZE, OFF_BD = OFF_WD + HID * NOUT, P = OFF_BD + NOUT;
    // bf16 copies: We^T, then per layer W1ab^T, W1ab, W2^T, W2, Wn1^T, Wn1, Wn2^T, Wn2
    static constexpr int B_WE = 0, B_L0 = B_WE + HID * NIN;
    static constexpr int B_W1T = 0, B_W1 = B_W1T + 128 * HID, B_W2T = B_W1 + HID * 128, B_W2 = B_W2T + HID * HID,
                         B_WN1T = B_W2 + HID * HID, B_WN1 = B_WN1T + HID * 128, B_WN2T = B_WN1 + 128 * HID,
                         B_WN2 = B_WN2T + HID * HID, B_LSIZE = B_WN2 + HID * HID;
    static constexpr int B_TOTAL = B_L0 + LAYERS * B_LSIZE;

    hipStream_t st = nullptr;
    float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    __bf16 *wb = nullptr;
    int adam_t = 0;
    // feature / target scales (set by the caller from its training data)
    float rho0 = 1.f, v_scale = 1.f, rho_scale = 1.f, nu_scale = 1.f, acc_scale = 1.f;
    std::vector<MpnSample> samples;

    explicit Mpn(uint64_t seed);
    ~Mpn();
    void set_params(const float *host);
    void get_params(float *host);
    void refresh_bf16();
    // activations for graphs up to n nodes
    void ensure(int n);
    // out [n][3] (device) for graph g with node features X0 (device)
    void forward(const MpnGraph &gr, const float *X0, float *out);
    // gradients of the mean squared error against y into g (zeroed first);
    // needs the activations of forward() on the same graph; loss on device
    void backward(const MpnGraph &gr, const float *X0, const float *y, float *d_loss);
    void adam(float lr);
    float *out_buf() { return out_; }

  private:
    int cap_ = 0;
    float *Zenc_ = nullptr, *H_[LAYERS + 1] = {}, *AB_[LAYERS] = {}, *S_[LAYERS] = {}, *M_[LAYERS] = {},
          *Z1_[LAYERS] = {}, *U_[LAYERS] = {}, *out_ = nullptr;
    float *dH_ = nullptr, *dT_ = nullptr, *dU_ = nullptr, *dZ1_ = nullptr, *dHM_ = nullptr, *dS_ = nullptr,
          *dAB_ = nullptr, *dZe_ = nullptr;
    void free_acts();
};

}  // namespace ssa
