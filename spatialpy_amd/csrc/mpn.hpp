// Host side of the learned surrogate (edge network for the SDPD pair
// accelerations): parameters, training samples, training and inference on
// the GPU (kernels in mpn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "engine.hpp"

namespace ssa {

// Feature and target scales: rho0, speed, density deviation, viscosity,
// pair acceleration, |v - vt|.
struct EcnScales {
    float rho0 = 1.f, vs = 1.f, rs = 1.f, ns = 1.f, acc = 1.f, ts = 1.f;
};

// Device view of one particle graph: per-particle records and the ELL
// lists (slot k of particle i at k * n + i).
struct EcnGraph {
    int n = 0, W = 0, dim = 3;
    double h = 1.0;
    const PRec *rec = nullptr;
    const int *deg = nullptr, *nidx = nullptr;
    const double *nr = nullptr;
};

// Materialised edges of a particle subset (a training sample).
struct EcnSample {
    int np = 0, ne = 0;
    float *X = nullptr;      // [ne][32] features
    float *Bv = nullptr;     // [ne][16] basis (12), envelope, pad
    int *pe = nullptr;       // [ne] owning particle (0..np-1)
    int *rowp = nullptr;     // [np + 1]
    float *y = nullptr;      // [np][3] targets / acc_scale
};

class Mpn {
  public:
    static constexpr int NF = 16, KIN = 32, HID = 64, NOUT = 4;
    // flat FP32 parameters: W1 [32][64], b1, W2 [64][64], b2, W3 [64][4], b3,
    // L [16][4] (linear skip from the features to the coefficients)
    static constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + KIN * HID, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID,
                         OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID * NOUT, OFF_L = OFF_B3 + NOUT,
                         P = OFF_L + NF * NOUT;
    // bf16 copies: W1t [64][32], W2t [64][64] (forward, out-major), W2
    // [64][64] (backward, in-major); the head W3 stays FP32
    static constexpr int B_W1T = 0, B_W2T = B_W1T + HID * KIN, B_W2 = B_W2T + HID * HID, B_TOTAL = B_W2 + HID * HID;

    hipStream_t st = nullptr;
    float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    __bf16 *wb = nullptr;
    int adam_t = 0;
    EcnScales sc;
    std::vector<EcnSample> samples;

    explicit Mpn(uint64_t seed);
    ~Mpn();
    void set_params(const float *host);
    void get_params(float *host);
    void refresh_bf16();

    // F (SoA [3][n]) = g + acc_scale * (predicted pair acceleration) for
    // moving particles, g for solid ones; on stream s.
    void infer(hipStream_t s, const EcnGraph &gr, const double *gvec, double *F, double cap = 0.0,
               unsigned long long *n_capped = nullptr);
    // materialise the edges of the n_sel particles sel_host[] of gr as a new
    // sample; targets from y_host ([n_sel][3], already / acc_scale) or, when
    // null, from F_dev (SoA [3][n] accelerations, gravity gvec) on the device
    int add_sample(hipStream_t s, const EcnGraph &gr, const int *deg_host, const int *sel_host, int n_sel,
                   const float *y_host, const double *F_dev, const double *gvec);
    void clear_samples();
    // training-path forward of sample k: predictions (acc_scale units) into
    // pred_ [np][3]; with loss != null also dL/dpred and the loss
    void forward(const EcnSample &s, float *loss);
    // gradients into g (zeroed first) after forward(s, loss)
    void backward(const EcnSample &s);
    void adam(float lr);
    // `iters` Adam steps: the even ones on the newest stored sample, the odd
    // ones on a sample drawn uniformly from all of them (seeded)
    void train_steps(int iters, float lr, uint64_t seed);
    // release the oldest stored sample
    void drop_oldest();
    const float *pred() const { return pred_; }

  private:
    size_t cap_e_ = 0, cap_p_ = 0;
    float *Z1_ = nullptr, *H1_ = nullptr, *Z2_ = nullptr, *H2_ = nullptr, *O_ = nullptr, *pred_ = nullptr,
          *D_ = nullptr, *dH_ = nullptr, *dZ_ = nullptr, *dloss_ = nullptr, *part_ = nullptr;
    void ensure(size_t ne, size_t np);
    void free_acts();
};

// launchers (mpn.hip)
// cap > 0: a moving particle's predicted pair acceleration is scaled down to
// magnitude cap where it is larger, and n_capped (if given) counts those
// particles
void ecn_infer(hipStream_t s, const EcnGraph &g, const EcnScales &sc, const __bf16 *wb, const float *w,
               const double *gr, double *F, double cap = 0.0, unsigned long long *n_capped = nullptr);
void ecn_gemm(hipStream_t s, int M, int K, const float *A, int lda, const __bf16 *Wt, int N, const float *bias, int act,
              float *Z, float *Y, int ldy);
// weight gradients are summed without atomics: per-block partials in
// `part` (ecn_wgrad_parts / 256 N / ecn_head_parts floats), then added in
// block order, so training is reproducible run to run
size_t ecn_wgrad_parts(int M, int K, int N);
void ecn_wgrad(hipStream_t s, int M, int K, const float *A, int lda, const float *dY, int ldy, int N, float *dW,
               float *part);
void ecn_colsum(hipStream_t s, int M, int N, const float *dY, int ldy, float *db, float *part);
void ecn_act_bwd(hipStream_t s, size_t m, const float *dY, const float *Z, float *dZ);
void ecn_edges(hipStream_t s, const EcnGraph &g, const EcnScales &sc, int np, const int *sel, const int *rowp,
               float *X, float *Bv, int *pe);
void ecn_target(hipStream_t s, int n, int np, const int *sel, const double *F, const double *g, float acc, float *y);
void ecn_pred(hipStream_t s, int np, const int *rowp, const float *O, const float *b3, const float *Bv, const float *y,
              float *pred, float *D, float *loss);
void ecn_head_fwd(hipStream_t s, int E, const float *H2, const float *X, const float *W3, const float *L, float *O);
// g_head: the gradient block W3 [64][4], b3 [4], L [16][4] (contiguous in
// the parameter layout)
size_t ecn_head_parts(int E);
void ecn_head_bwd(hipStream_t s, int E, const int *pe, const float *Bv, const float *D, const float *H2,
                  const float *Z2, const float *X, const float *W3, float *dZ2, float *g_head, float *part);
void ecn_adam(hipStream_t s, int m, float *w, const float *g, float *mo, float *ve, float lr, int t);
void ecn_bf16(hipStream_t s, const float *w, __bf16 *wb);
// over moving particles: acc[0] += |Fref - Fnn|^2, acc[1] += |Fref - g|^2,
// acc[2] += 1
void ecn_err(hipStream_t s, int n, const double *Fref, const double *Fnn, const int *solid, const double *g,
             double *acc);

}  // namespace ssa
