// Host side of the learned surrogate (see mpn.hip for the network).
#include <hip/hip_runtime.h>

#include <algorithm>
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
    // He-normal hidden weights (W1 rows past the 16 features stay zero),
    // small biases, a small head
    std::vector<float> h(P, 0.f);
    std::mt19937_64 rng(seed);
    auto normal = [&](int off, int count, float sd) {
        std::normal_distribution<float> d(0.f, sd);
        for (int i = 0; i < count; i++) h[off + i] = d(rng);
    };
    normal(OFF_W1, NF * HID, std::sqrt(2.f / NF));
    normal(OFF_B1, HID, 0.1f);
    normal(OFF_W2, HID * HID, std::sqrt(2.f / HID));
    normal(OFF_B2, HID, 0.1f);
    normal(OFF_W3, HID * NOUT, 0.1f / std::sqrt((float)HID));
    normal(OFF_L, NF * NOUT, 0.1f / std::sqrt((float)NF));
    set_params(h.data());
}

Mpn::~Mpn() {
    free_acts();
    clear_samples();
    release(dloss_);
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

void Mpn::refresh_bf16() { ecn_bf16(st, w, wb); }

void Mpn::infer(hipStream_t s, const EcnGraph &gr, const double *gvec, double *F, double cap,
                unsigned long long *n_capped) {
    // the weights were last written on the network's own stream
    HIP_OK(hipStreamSynchronize(st));
    ecn_infer(s, gr, sc, wb, w, gvec, F, cap, n_capped);
}

void Mpn::clear_samples() {
    for (auto &s : samples) {
        release(s.X); release(s.Bv); release(s.pe); release(s.rowp); release(s.y);
    }
    samples.clear();
}

int Mpn::add_sample(hipStream_t s, const EcnGraph &gr, const int *deg_host, const int *sel_host, int n_sel,
                    const float *y_host, const double *F_dev, const double *gvec) {
    std::vector<int> rowp(n_sel + 1, 0);
    for (int p = 0; p < n_sel; p++) rowp[p + 1] = rowp[p] + deg_host[sel_host[p]];
    EcnSample smp;
    smp.np = n_sel;
    smp.ne = rowp[n_sel];
    smp.X = alloc<float>((size_t)smp.ne * KIN);
    smp.Bv = alloc<float>((size_t)smp.ne * 16);
    smp.pe = alloc<int>(smp.ne);
    smp.rowp = alloc<int>(n_sel + 1);
    smp.y = alloc<float>((size_t)n_sel * 3);
    int *sel = alloc<int>(n_sel);
    HIP_OK(hipMemcpyAsync(sel, sel_host, sizeof(int) * n_sel, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(smp.rowp, rowp.data(), sizeof(int) * (n_sel + 1), hipMemcpyHostToDevice, s));
    if (y_host)
        HIP_OK(hipMemcpyAsync(smp.y, y_host, sizeof(float) * 3 * n_sel, hipMemcpyHostToDevice, s));
    else
        ecn_target(s, gr.n, n_sel, sel, F_dev, gvec, sc.acc, smp.y);
    ecn_edges(s, gr, sc, n_sel, sel, smp.rowp, smp.X, smp.Bv, smp.pe);
    HIP_OK(hipStreamSynchronize(s));
    release(sel);
    samples.push_back(smp);
    return (int)samples.size() - 1;
}

void Mpn::free_acts() {
    release(Z1_); release(H1_); release(Z2_); release(H2_); release(O_); release(pred_); release(D_);
    release(dH_); release(dZ_); release(part_);
    cap_e_ = cap_p_ = 0;
}

void Mpn::ensure(size_t ne, size_t np) {
    if (ne <= cap_e_ && np <= cap_p_) return;
    free_acts();
    cap_e_ = ne; cap_p_ = np;
    Z1_ = alloc<float>(ne * HID); H1_ = alloc<float>(ne * HID); Z2_ = alloc<float>(ne * HID);
    H2_ = alloc<float>(ne * HID); O_ = alloc<float>(ne * NOUT);
    dH_ = alloc<float>(ne * HID); dZ_ = alloc<float>(ne * HID);
    pred_ = alloc<float>(np * 3); D_ = alloc<float>(np * 3);
    const size_t parts = std::max({ecn_wgrad_parts((int)ne, HID, HID), (size_t)256 * HID, ecn_head_parts((int)ne)});
    part_ = alloc<float>(parts);
}

void Mpn::forward(const EcnSample &s, float *loss) {
    ensure((size_t)s.ne, (size_t)s.np);
    const int E = s.ne;
    ecn_gemm(st, E, KIN, s.X, KIN, wb + B_W1T, HID, w + OFF_B1, 1, Z1_, H1_, HID);
    ecn_gemm(st, E, HID, H1_, HID, wb + B_W2T, HID, w + OFF_B2, 1, Z2_, H2_, HID);
    // FP32 head; k_ecn_pred adds b3
    ecn_head_fwd(st, E, H2_, s.X, w + OFF_W3, w + OFF_L, O_);
    ecn_pred(st, s.np, s.rowp, O_, w + OFF_B3, s.Bv, loss ? s.y : nullptr, pred_, D_, loss);
}

void Mpn::backward(const EcnSample &s) {
    const int E = s.ne;
    HIP_OK(hipMemsetAsync(g, 0, sizeof(float) * P, st));
    // head: dW3, db3, dL and dZ2 = (dO W3^T) hs'(Z2)
    static_assert(OFF_B3 == OFF_W3 + HID * NOUT && OFF_L == OFF_B3 + NOUT, "head gradient block is contiguous");
    ecn_head_bwd(st, E, s.pe, s.Bv, D_, H2_, Z2_, s.X, w + OFF_W3, dZ_, g + OFF_W3, part_);
    ecn_wgrad(st, E, HID, H1_, HID, dZ_, HID, HID, g + OFF_W2, part_);
    ecn_colsum(st, E, HID, dZ_, HID, g + OFF_B2, part_);
    // dH1 = dZ2 W2^T, dZ1 = dH1 hs'(Z1)
    ecn_gemm(st, E, HID, dZ_, HID, wb + B_W2, HID, nullptr, 0, nullptr, dH_, HID);
    ecn_act_bwd(st, (size_t)E * HID, dH_, Z1_, dZ_);
    ecn_wgrad(st, E, KIN, s.X, KIN, dZ_, HID, HID, g + OFF_W1, part_);
    ecn_colsum(st, E, HID, dZ_, HID, g + OFF_B1, part_);
}

void Mpn::train_steps(int iters, float lr, uint64_t seed) {
    if (samples.empty() || iters < 1) return;
    if (!dloss_) dloss_ = alloc<float>(1);
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<size_t> pick(0, samples.size() - 1);
    for (int it = 0; it < iters; it++) {
        // every other step on the newest sample, the rest drawn from all
        const EcnSample &s = (it % 2 == 0) ? samples.back() : samples[pick(rng)];
        forward(s, dloss_);
        backward(s);
        adam(lr);
    }
    HIP_OK(hipStreamSynchronize(st));
}

void Mpn::drop_oldest() {
    if (samples.empty()) return;
    EcnSample &s = samples.front();
    release(s.X); release(s.Bv); release(s.pe); release(s.rowp); release(s.y);
    samples.erase(samples.begin());
}

void Mpn::adam(float lr) {
    adam_t++;
    ecn_adam(st, P, w, g, m1, m2, lr, adam_t);
    refresh_bf16();
}

}  // namespace ssa
