// Host side of the message-passing surrogate (see mpn.hip for the network).
#include <hip/hip_runtime.h>

#include <cmath>
#include <random>
#include <vector>

#include "mpn.hpp"

namespace ssa {

template <class T> static T *alloc(size_t count) {
    void *p = nullptr;
    HIP_OK(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
    return (T *)p;
}
template <class T> static void release(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

Mpn::Mpn(uint64_t seed) {
    HIP_OK(hipStreamCreate(&st));
    w = alloc<float>(P); g = alloc<float>(P); m1 = alloc<float>(P); m2 = alloc<float>(P);
    wb = alloc<__bf16>(B_TOTAL);
    HIP_OK(hipMemsetAsync(m1, 0, sizeof(float) * P, st));
    HIP_OK(hipMemsetAsync(m2, 0, sizeof(float) * P, st));
    // Glorot-uniform weights, zero biases, a small decoder
    std::vector<float> h(P, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_in, int fan_out, int rows, float gain) {
        const float a = gain * std::sqrt(6.0f / (float)(fan_in + fan_out));
        std::uniform_real_distribution<float> u(-a, a);
        for (int i = 0; i < rows * fan_out; i++) h[off + i] = u(rng);
    };
    fill(OFF_WE, 10, HID, 10, 1.f);  // rows past the 10 features stay 0
    for (int l = 0; l < LAYERS; l++) {
        const int b = L0 + l * LSIZE;
        {   // W1ab [64][128] (A block | B block) and W1c [4][64]: the edge MLP's first layer, fan-in 132
            const float a = std::sqrt(6.0f / (float)(2 * HID + 4 + HID));
            std::uniform_real_distribution<float> u(-a, a);
            for (int i = 0; i < HID * 128; i++) h[b + W1AB + i] = u(rng);
            for (int i = 0; i < 4 * HID; i++) h[b + W1C + i] = u(rng);
        }
        fill(b + W2, HID, HID, HID, 1.f);
        fill(b + WN1, 2 * HID, HID, 2 * HID, 1.f);
        fill(b + WN2, HID, HID, HID, 0.5f);
    }
    fill(OFF_WD, HID, NOUT, HID, 0.1f);
    set_params(h.data());
}

Mpn::~Mpn() {
    free_acts();
    for (auto &s : samples) {
        release(s.x); release(s.X0); release(s.deg); release(s.nidx); release(s.y);
    }
    release(w); release(g); release(m1); release(m2); release(wb);
    if (st) (void)hipStreamDestroy(st);
}

void Mpn::set_params(const float *host) {
    HIP_OK(hipMemcpyAsync(w, host, sizeof(float) * P, hipMemcpyHostToDevice, st));
    refresh_bf16();
    HIP_OK(hipStreamSynchronize(st));
}
void Mpn::get_params(float *host) {
    HIP_OK(hipMemcpyAsync(host, w, sizeof(float) * P, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
}

void Mpn::refresh_bf16() {
    mpn_to_bf16_t(st, NIN, NIN, HID, w + OFF_WE, wb + B_WE);
    for (int l = 0; l < LAYERS; l++) {
        const float *p = w + L0 + l * LSIZE;
        __bf16 *q = wb + B_L0 + l * B_LSIZE;
        mpn_to_bf16_t(st, HID, HID, 128, p + W1AB, q + B_W1T);
        mpn_to_bf16(st, HID, 128, p + W1AB, q + B_W1);
        mpn_to_bf16_t(st, HID, HID, HID, p + W2, q + B_W2T);
        mpn_to_bf16(st, HID, HID, p + W2, q + B_W2);
        mpn_to_bf16_t(st, 128, 128, HID, p + WN1, q + B_WN1T);
        mpn_to_bf16(st, 128, HID, p + WN1, q + B_WN1);
        mpn_to_bf16_t(st, HID, HID, HID, p + WN2, q + B_WN2T);
        mpn_to_bf16(st, HID, HID, p + WN2, q + B_WN2);
    }
}

void Mpn::free_acts() {
    release(Zenc_);
    for (int l = 0; l <= LAYERS; l++) release(H_[l]);
    for (int l = 0; l < LAYERS; l++) { release(AB_[l]); release(S_[l]); release(M_[l]); release(Z1_[l]); release(U_[l]); }
    release(out_); release(dH_); release(dT_); release(dU_); release(dZ1_); release(dHM_); release(dS_); release(dAB_);
    release(dZe_);
    cap_ = 0;
}

void Mpn::ensure(int n) {
    if (n <= cap_) return;
    free_acts();
    const size_t m = (size_t)n;
    Zenc_ = alloc<float>(m * HID);
    for (int l = 0; l <= LAYERS; l++) H_[l] = alloc<float>(m * HID);
    for (int l = 0; l < LAYERS; l++) {
        AB_[l] = alloc<float>(m * 128); S_[l] = alloc<float>(m * HID); M_[l] = alloc<float>(m * HID);
        Z1_[l] = alloc<float>(m * HID); U_[l] = alloc<float>(m * HID);
    }
    out_ = alloc<float>(m * NOUT);
    dH_ = alloc<float>(m * HID); dT_ = alloc<float>(m * HID); dU_ = alloc<float>(m * HID);
    dZ1_ = alloc<float>(m * HID); dHM_ = alloc<float>(m * 128); dS_ = alloc<float>(m * HID);
    dAB_ = alloc<float>(m * 128); dZe_ = alloc<float>(m * HID);
    cap_ = n;
}

void Mpn::forward(const MpnGraph &gr, const float *X0, float *out) {
    const int n = gr.n;
    ensure(n);
    mpn_gemm(st, n, NIN, NIN, X0, NIN, nullptr, 0, wb + B_WE, HID, w + OFF_BE, nullptr, 1, nullptr, Zenc_, H_[0]);
    for (int l = 0; l < LAYERS; l++) {
        const float *p = w + L0 + l * LSIZE;
        const __bf16 *q = wb + B_L0 + l * B_LSIZE;
        mpn_gemm(st, n, HID, HID, H_[l], HID, nullptr, 0, q + B_W1T, 128, nullptr, nullptr, 0, nullptr, nullptr, AB_[l]);
        mpn_edge_fwd(st, gr, AB_[l], p + W1C, p + B1, S_[l]);
        mpn_gemm(st, n, HID, HID, S_[l], HID, nullptr, 0, q + B_W2T, HID, p + B2, gr.deg, 0, nullptr, nullptr, M_[l]);
        mpn_gemm(st, n, 128, HID, H_[l], HID, M_[l], HID, q + B_WN1T, HID, p + BN1, nullptr, 1, nullptr, Z1_[l], U_[l]);
        mpn_gemm(st, n, HID, HID, U_[l], HID, nullptr, 0, q + B_WN2T, HID, p + BN2, nullptr, 0, H_[l], nullptr,
                 H_[l + 1]);
    }
    mpn_decode(st, n, H_[LAYERS], w + OFF_WD, w + OFF_BD, out ? out : out_);
}

void Mpn::backward(const MpnGraph &gr, const float *X0, const float *y, float *d_loss) {
    const int n = gr.n;
    const size_t nh = (size_t)n * HID;
    HIP_OK(hipMemsetAsync(g, 0, sizeof(float) * P, st));
    mpn_decode_bwd(st, n, H_[LAYERS], w + OFF_WD, out_, y, dH_, g + OFF_WD, g + OFF_BD, d_loss);
    for (int l = LAYERS - 1; l >= 0; l--) {
        const float *p = w + L0 + l * LSIZE;
        float *gp = g + L0 + l * LSIZE;
        const __bf16 *q = wb + B_L0 + l * B_LSIZE;
        // H' = H + U Wn2 + bn2
        mpn_gemm(st, n, HID, HID, dH_, HID, nullptr, 0, q + B_WN2, HID, nullptr, nullptr, 0, nullptr, nullptr, dU_);
        mpn_wgrad(st, n, HID, HID, U_[l], HID, nullptr, 0, dH_, HID, HID, gp + WN2);
        mpn_colsum(st, n, HID, dH_, HID, nullptr, gp + BN2);
        // U = silu([H|M] Wn1 + bn1)
        mpn_silu_bwd(st, nh, dU_, Z1_[l], dZ1_);
        mpn_gemm(st, n, HID, HID, dZ1_, HID, nullptr, 0, q + B_WN1, 128, nullptr, nullptr, 0, nullptr, nullptr, dHM_);
        mpn_wgrad(st, n, 128, HID, H_[l], HID, M_[l], HID, dZ1_, HID, HID, gp + WN1);
        mpn_colsum(st, n, HID, dZ1_, HID, nullptr, gp + BN1);
        // M = S W2 + deg b2   (dM = dHM[:, 64:128])
        mpn_gemm(st, n, HID, HID, dHM_ + HID, 128, nullptr, 0, q + B_W2, HID, nullptr, nullptr, 0, nullptr, nullptr, dS_);
        mpn_wgrad(st, n, HID, HID, S_[l], HID, nullptr, 0, dHM_ + HID, 128, HID, gp + W2);
        mpn_colsum(st, n, HID, dHM_ + HID, 128, gr.deg, gp + B2);
        // S = sum_j silu(A_i + B_j + e W1c + b1)
        HIP_OK(hipMemsetAsync(dAB_, 0, sizeof(float) * (size_t)n * 128, st));
        mpn_edge_bwd(st, gr, AB_[l], p + W1C, p + B1, dS_, dAB_, gp + W1C, gp + B1);
        // AB = H W1ab
        mpn_wgrad(st, n, HID, HID, H_[l], HID, nullptr, 0, dAB_, 128, 128, gp + W1AB);
        // dH_in = dH (residual) + dHM[:, 0:64] + dAB W1ab^T
        mpn_add(st, n, HID, dH_, dHM_, 128);
        mpn_gemm(st, n, 128, 128, dAB_, 128, nullptr, 0, q + B_W1, HID, nullptr, nullptr, 0, dH_, nullptr, dT_);
        std::swap(dH_, dT_);
    }
    // H0 = silu(X0 We + be)
    mpn_silu_bwd(st, nh, dH_, Zenc_, dZe_);
    mpn_wgrad(st, n, NIN, NIN, X0, NIN, nullptr, 0, dZe_, HID, HID, g + OFF_WE);
    mpn_colsum(st, n, HID, dZe_, HID, nullptr, g + OFF_BE);
}

void Mpn::adam(float lr) {
    adam_t++;
    mpn_adam(st, P, w, g, m1, m2, lr, adam_t);
    refresh_bf16();
}

}  // namespace ssa
