// Learned surrogate for the SDPD accelerations on gfx950: an edge network
// that keeps the pair structure of the SDPD force sum.
//
// Per neighbour pair (i, j) of the fixed-radius lists:
//   f_ij   16 rotation-invariant features (padded to 32)
//   o_ij   = W3 hs(W2 hs(W1 f + b1) + b2) + L f + b3    32 -> 64 -> 64 -> 4,
//            with a linear skip L from the 16 features
//   m_ij   = u(R) (o0 dx/h + o1 dv/v_s + o2 v_i/v_s + o3 v_j/v_s)
//   a_i    = acc_scale * sum_j m_ij                     (the pair acceleration)
// u(R) = (1-R)^2 R / (R + 0.001), R = r/h, is a fixed cut-off envelope and
// hs(z) = z min(max(z + 3, 0), 6) / 6 the hidden activation.  The SDPD pair
// term is a dx + b dv + c v_i + d v_j with invariant coefficients
// (pressure, viscosity, transport velocity), so the network only has to
// learn four scalar functions; the prediction rotates with the frame.
//
// Products: the two hidden layers on v_mfma_f32_16x16x32_bf16 with bf16
// weights and FP32 accumulation.  Activations enter as two bf16 halves
// (hi = bf16(a), lo = bf16(a - hi)), two products per tile, so they carry
// ~16 significant bits; the weights are trained as the bf16 values the
// products use.  The head (64 -> 4), whose weights scale the four
// coefficients directly, stays FP32 on the vector units.
//
// Inference (k_ecn_infer) is one fused kernel: a wavefront owns 16
// particles and walks their list slots together; each step is a 16-edge
// tile pushed through the hidden layers in transposed form (weights are the
// A operand, edges the 16 columns), so the four lanes of an edge column
// hold 16 of its 64 hidden values and a layer's output registers are the
// next layer's B operand as they stand (the k order inside a product is
// permuted to match, and the weight rows are read in that order); the head
// sums each lane's 16 products and reduces over the four lanes.  Nothing
// but the final accelerations is written.
//
// Training materialises the edges of a particle subset ([E][32] features,
// [E][16] basis vectors and envelope) and runs the layers as row-tiled
// GEMMs (k_gemm / k_gemm_wgrad) with the head, loss and their backward in
// small edge- and particle-parallel kernels; Adam on FP32 master weights.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mpn.hpp"

namespace ssa {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float hsw(float z) { return z * fminf(fmaxf(z + 3.f, 0.f), 6.f) * (1.f / 6.f); }
__device__ __forceinline__ float hsw_grad(float z) {
    return z < -3.f ? 0.f : (z > 3.f ? 1.f : (2.f * z + 3.f) * (1.f / 6.f));
}

__device__ __forceinline__ void split8(const float v[8], bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const __bf16 h = (__bf16)v[q];
        hi[q] = h;
        lo[q] = (__bf16)(v[q] - (float)h);
    }
}

// ---------------------------------------------------------------------------
// Edge features.  Particle i's record is `pi`, neighbour j's is read from
// the records; rd is the listed distance.  Returns false (and zeros) for
// the slots the pair loop skips (rd/h > 1 or rd == 0).
struct EdgeOut {
    float f[Mpn::NF];
    float b[12];         // dx/h, dv/v_s, v_i/v_s, v_j/v_s
    float u;
};

__device__ __forceinline__ void rec_load(const PRec *r, double x[3], double v[3], double vt[3], double &nu,
                                         double &rho, double &mass, int &solid) {
    const double2 *p = (const double2 *)r;
    const double2 a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3], a4 = p[4], a5 = p[5];
    const int4 t = ((const int4 *)p)[6];
    x[0] = a0.x; x[1] = a0.y; x[2] = a1.x;
    v[0] = a1.y; v[1] = a2.x; v[2] = a2.y;
    vt[0] = a3.x; vt[1] = a3.y; vt[2] = a4.x;
    nu = a4.y; rho = a5.x; mass = a5.y;
    solid = t.y;
}

struct PState {
    double x[3], v[3], vt[3], nu, rho, mass;
    int solid;
};

// Scales as reciprocals (the features multiply by them).
struct EcnInv {
    double hinv, rho0inv_d;
    float rho0inv, vsinv, rsinv, nsinv, tsinv;
};
__host__ __device__ inline EcnInv ecn_inv(const EcnGraph &g, const EcnScales &s) {
    return EcnInv{1.0 / g.h, 1.0 / (double)s.rho0, 1.f / s.rho0, 1.f / s.vs, 1.f / s.rs, 1.f / s.ns, 1.f / s.ts};
}

// Particle i's side of the features, computed once per particle.
struct PSide {
    double x[3], v[3], r_d;     // position, velocity, rho/rho0 (FP64)
    float vs[3], tv[3];         // v/v_s, v - vt
    float r, rinv, pg, minv, nu, solid;
};
__device__ __forceinline__ PSide pside(const PState &p, const EcnInv &iv) {
    PSide o;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        o.x[c] = p.x[c];
        o.v[c] = p.v[c];
        o.vs[c] = (float)p.v[c] * iv.vsinv;
        o.tv[c] = (float)(p.v[c] - p.vt[c]);
    }
    o.r_d = p.rho * iv.rho0inv_d;
    o.r = (float)o.r_d;
    o.rinv = 1.f / o.r;
    o.pg = (o.r - 1.f) * o.rinv * o.rinv;
    o.minv = 1.f / (float)p.mass;
    o.nu = (float)p.nu * iv.nsinv;
    o.solid = p.solid ? 1.f : 0.f;
    return o;
}

// Features (FP64 differences, the rest in FP32; r = rho/rho0, n = m/m_i):
//   0 R = r/h            1 (1/(r_i r_j) - 1)/rho_s: the viscous term's
//   density factor       2, 3 (r - 1)/rho_s of i, j
//   4 sigma: the pair loop's pressure sign rule taken (P_i/rho_i^2 +
//     P_j/rho_j^2 < 0)   5 dv.dx/(h v_s)   6, 7 (v - vt).dx/(h tv_s) of i, j
//   8 (s P_i/rho_i^2 + P_j/rho_j^2) in units of P0/rho0^2 rho_s, s = -1
//     when sigma, else 1 (the pressure factor of the pair term)
//   9 m_j/m_i - 1        10, 11 nu/nu_s of i, j    12, 13 solid flags
//   14, 15 features 6, 7 times r_i resp. r_j times V = (1/r_i^2 + n_j^2/r_j^2)/2:
//     the transport tensor's factors ((m_i/rho_i)^2 + (m_j/rho_j)^2)/m_i and
//     rho_i resp. rho_j (model.cpp:127-130)
// The network gets the pair loop's scalar factors; it learns their
// constants and how they combine (including where masses, viscosities or
// solid neighbours vary).
__device__ __forceinline__ bool ecn_edge(const EcnGraph &g, const EcnInv &iv, const PSide &pi, bool inlist, int j,
                                         double rd, EdgeOut &o) {
    const bool valid = inlist && !(rd > g.h) && rd != 0.0;
    PState pj;
    rec_load(g.rec + j, pj.x, pj.v, pj.vt, pj.nu, pj.rho, pj.mass, pj.solid);
    float d[3], dv[3], vj[3];
    float dd = 0.f, ti = 0.f, tj = 0.f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        d[c] = (float)((pi.x[c] - pj.x[c]) * iv.hinv);
        dv[c] = (float)(pi.v[c] - pj.v[c]) * iv.vsinv;
        vj[c] = (float)pj.v[c] * iv.vsinv;
        dd += dv[c] * d[c];
        ti += pi.tv[c] * d[c];
        tj += (float)(pj.v[c] - pj.vt[c]) * d[c];
    }
    const float R = (float)(rd * iv.hinv), q = 1.f - R;
    const double rj_d = pj.rho * iv.rho0inv_d;
    // sign of d_i/r_i^2 + d_j/r_j^2 (r: rho/rho0, d = r - 1) without division, in FP64
    const bool sigma = (pi.r_d - 1.0) * rj_d * rj_d + (rj_d - 1.0) * pi.r_d * pi.r_d < 0.0;
    const float rj = (float)rj_d;
    const float rjinv = __builtin_amdgcn_rcpf(rj);
    const float pgj = (rj - 1.f) * rjinv * rjinv;
    const float nj = (float)pj.mass * pi.minv;
    const float V = 0.5f * (pi.rinv * pi.rinv + nj * nj * rjinv * rjinv);
    const float tis = ti * iv.tsinv, tjs = tj * iv.tsinv;
    const float f[Mpn::NF] = {R, (pi.rinv * rjinv - 1.f) * iv.rsinv, (pi.r - 1.f) * iv.rsinv, (rj - 1.f) * iv.rsinv,
                              sigma ? 1.f : 0.f, dd, tis, tjs, ((sigma ? -pi.pg : pi.pg) + pgj) * iv.rsinv,
                              nj - 1.f, pi.nu, (float)pj.nu * iv.nsinv, pi.solid, pj.solid ? 1.f : 0.f,
                              tis * pi.r * V, tjs * rj * V};
#pragma unroll
    for (int k = 0; k < Mpn::NF; k++) o.f[k] = valid ? f[k] : 0.f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        o.b[c] = valid ? d[c] : 0.f;
        o.b[3 + c] = valid ? dv[c] : 0.f;
        o.b[6 + c] = valid ? pi.vs[c] : 0.f;
        o.b[9 + c] = valid ? vj[c] : 0.f;
    }
    o.u = valid ? __fdividef(q * q * R, R + 0.001f) : 0.f;
    return valid;
}

__device__ __forceinline__ PState load_state(const EcnGraph &g, int i) {
    PState p;
    rec_load(g.rec + i, p.x, p.v, p.vt, p.nu, p.rho, p.mass, p.solid);
    return p;
}

// ---------------------------------------------------------------------------
// Fused inference.  256 threads = 4 wavefronts x 16 particles.  Weights:
// W1t [64][32], W2t [64][64] bf16 (out-major) and the FP32 head W3 [64][4],
// staged in LDS.
// F (SoA [3][n]) = g + acc_scale * a_i for moving particles, g for solid
// ones (and for components past dim).
constexpr int ECN_PER_WAVE = 16;

__global__ void __launch_bounds__(256) k_ecn_infer(EcnGraph g, EcnScales s, const __bf16 *__restrict__ wb,
                                                   const float *__restrict__ w, double g0, double g1, double g2,
                                                   double *__restrict__ F, double cap,
                                                   unsigned long long *__restrict__ n_capped) {
    __shared__ __align__(16) __bf16 sw1[Mpn::HID * Mpn::KIN];
    __shared__ __align__(16) __bf16 sw2[Mpn::HID * Mpn::HID];
    __shared__ __align__(16) float sw3f[Mpn::HID * Mpn::NOUT];
    __shared__ __align__(16) float sb1[Mpn::HID], sb2[Mpn::HID], sb3[Mpn::NOUT];
    for (int q = threadIdx.x; q < Mpn::HID; q += blockDim.x) {
        sb1[q] = w[Mpn::OFF_B1 + q];
        sb2[q] = w[Mpn::OFF_B2 + q];
        if (q < Mpn::NOUT) sb3[q] = w[Mpn::OFF_B3 + q];
    }
    for (int q = threadIdx.x; q < Mpn::HID * Mpn::NOUT; q += blockDim.x) sw3f[q] = w[Mpn::OFF_W3 + q];
    __shared__ __align__(16) float slf[Mpn::NF * Mpn::NOUT];
    for (int q = threadIdx.x; q < Mpn::NF * Mpn::NOUT; q += blockDim.x) slf[q] = w[Mpn::OFF_L + q];
    for (int q = threadIdx.x; q < Mpn::HID * Mpn::KIN / 8; q += blockDim.x)
        ((bf16x8 *)sw1)[q] = ((const bf16x8 *)(wb + Mpn::B_W1T))[q];
    for (int q = threadIdx.x; q < Mpn::HID * Mpn::HID / 8; q += blockDim.x)
        ((bf16x8 *)sw2)[q] = ((const bf16x8 *)(wb + Mpn::B_W2T))[q];
    __syncthreads();
    const int nb = (g.n + 4 * ECN_PER_WAVE - 1) / (4 * ECN_PER_WAVE);
    const int chunk = xcd_chunk(nb);
    if (chunk < 0) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int e = lane & 15, grp = lane >> 4;
    const int p = (chunk * 4 + wave) * ECN_PER_WAVE + e;
    const bool live = p < g.n;
    const int ps = live ? p : g.n - 1;
    const EcnInv iv = ecn_inv(g, s);
    const PSide pi = pside(load_state(g, ps), iv);
    const int dg = live ? g.deg[ps] : 0;
    int maxdeg = dg;
    for (int o = 32; o > 0; o >>= 1) maxdeg = max(maxdeg, __shfl_xor(maxdeg, o, 64));

    // this lane's biases (read from LDS at each use): hidden units
    // 16t + 4 grp + r, outputs 4 grp + r
    const float4 *b1v = (const float4 *)sb1 + grp, *b2v = (const float4 *)sb2 + grp;

    float acc[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < maxdeg; k++) {
        const bool inlist = k < dg;
        const size_t slot = (size_t)k * g.n + ps;
        const int j = inlist ? g.nidx[slot] : ps;
        const double rd = inlist ? g.nr[slot] : 0.0;
        EdgeOut eo;
        ecn_edge(g, iv, pi, inlist, j, rd, eo);
        // layer 1: B operand = features 8 grp .. 8 grp + 7 of edge e
        float xv[8];
#pragma unroll
        for (int q = 0; q < 8; q++) xv[q] = grp == 0 ? eo.f[q] : (grp == 1 ? eo.f[8 + q] : 0.f);
        bf16x8 xh, xl;
        split8(xv, xh, xl);
        f32x4 d1[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const bf16x8 a = *(const bf16x8 *)(sw1 + (16 * t + e) * Mpn::KIN + 8 * grp);
            d1[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xh, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            d1[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xl, d1[t], 0, 0, 0);
        }
        // layer 2: chunk c's k-slot 8 grp + q is hidden unit 32c + 16 (q>>2) + 4 grp + (q&3)
        bf16x8 hh[2], hl[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            float hv[8];
            const float4 ba = b1v[8 * c], bb = b1v[8 * c + 4];
            const float bav[4] = {ba.x, ba.y, ba.z, ba.w}, bbv[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
            for (int r = 0; r < 4; r++) {
                hv[r] = hsw(d1[2 * c][r] + bav[r]);
                hv[4 + r] = hsw(d1[2 * c + 1][r] + bbv[r]);
            }
            split8(hv, hh[c], hl[c]);
        }
        f32x4 d2[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            d2[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const __bf16 *row = sw2 + (16 * t + e) * Mpn::HID + 32 * c + 4 * grp;
                const bf16x4 lo4 = *(const bf16x4 *)row, hi4 = *(const bf16x4 *)(row + 16);
                const bf16x8 a = bf16x8{lo4[0], lo4[1], lo4[2], lo4[3], hi4[0], hi4[1], hi4[2], hi4[3]};
                d2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hh[c], d2[t], 0, 0, 0);
                d2[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hl[c], d2[t], 0, 0, 0);
            }
        }
        // head in FP32 (64 -> 4): each lane's 16 hidden values against the
        // FP32 weights, then the sum over the edge's four lanes
        float o[4] = {0.f, 0.f, 0.f, 0.f};
        // an offset the compiler cannot see through keeps the head's 64
        // weights per lane in LDS instead of registers across the loop
        int hoff = 4 * grp;
        asm volatile("" : "+v"(hoff));
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const float4 bt = b2v[4 * t];
            const float btv[4] = {bt.x, bt.y, bt.z, bt.w};
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float hv = hsw(d2[t][r] + btv[r]);
                const float4 wr = ((const float4 *)sw3f)[16 * t + hoff + r];
                o[0] += hv * wr.x; o[1] += hv * wr.y; o[2] += hv * wr.z; o[3] += hv * wr.w;
            }
        }
        // the linear skip: lane group grp adds features 4 grp .. 4 grp + 3
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float fv = grp == 0 ? eo.f[r] : grp == 1 ? eo.f[4 + r] : grp == 2 ? eo.f[8 + r] : eo.f[12 + r];
            const float4 lr = ((const float4 *)slf)[hoff + r];
            o[0] += fv * lr.x; o[1] += fv * lr.y; o[2] += fv * lr.z; o[3] += fv * lr.w;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            o[k] += __shfl_xor(o[k], 16, 64);
            o[k] += __shfl_xor(o[k], 32, 64);
        }
        const float4 b3v = *(const float4 *)sb3;
        const float o0 = o[0] + b3v.x, o1 = o[1] + b3v.y, o2 = o[2] + b3v.z, o3 = o[3] + b3v.w;
#pragma unroll
        for (int c = 0; c < 3; c++)
            acc[c] += eo.u * (o0 * eo.b[c] + o1 * eo.b[3 + c] + o2 * eo.b[6 + c] + o3 * eo.b[9 + c]);
    }
    if (grp == 0 && live) {
        const bool mv = pi.solid == 0.f;
        const double gg[3] = {g0, g1, g2};
        double a[3];
        double a2 = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            a[c] = (mv && c < g.dim) ? (double)(s.acc * acc[c]) : 0.0;
            a2 += a[c] * a[c];
        }
        double f = 1.0;
        if (cap > 0.0 && a2 > cap * cap) {
            f = cap / sqrt(a2);
            if (n_capped) atomicAdd(n_capped, 1ULL);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) F[(size_t)c * g.n + p] = gg[c] + a[c] * f;
    }
}

// ---------------------------------------------------------------------------
// Training path.
//
// Y[M][N] = act(A[M][K] W[K][N] + bias) with Wt: bf16 [N][K] (the weights
// transposed); Z (optional) receives the pre-activation.  A is FP32 and
// enters as bf16 hi + lo halves.  K % 32 == 0, N % 16 == 0, N <= 64, K <= 64.
// One wavefront computes 16 rows x all N columns.
template <int N>
__global__ void __launch_bounds__(256) k_gemm(int M, int K, const float *__restrict__ A, int lda,
                                              const __bf16 *__restrict__ Wt, const float *__restrict__ bias, int act,
                                              float *__restrict__ Z, float *__restrict__ Y, int ldy) {
    __shared__ __align__(16) __bf16 wsh[N * 64];
    for (int i = threadIdx.x; i < N * K; i += blockDim.x) wsh[i] = Wt[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * 4 + wave) * 16;
    if (r0 >= M) return;
    const int row = r0 + (lane & 15);
    const int kq = (lane >> 4) * 8;
    constexpr int NT = N / 16;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        const int kc = k0 + kq;
        float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (row < M) {
            const float4 lo = *(const float4 *)(A + (size_t)row * lda + kc);
            const float4 hi = *(const float4 *)(A + (size_t)row * lda + kc + 4);
            v[0] = lo.x; v[1] = lo.y; v[2] = lo.z; v[3] = lo.w; v[4] = hi.x; v[5] = hi.y; v[6] = hi.z; v[7] = hi.w;
        }
        bf16x8 ah, al;
        split8(v, ah, al);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int col = t * 16 + (lane & 15);
            const bf16x8 b = *(const bf16x8 *)(wsh + col * K + kc);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, b, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int col = t * 16 + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = r0 + (lane >> 4) * 4 + i;
            if (r >= M) continue;
            float z = acc[t][i];
            if (bias) z += bias[col];
            if (Z) Z[(size_t)r * N + col] = z;
            Y[(size_t)r * ldy + col] = act ? hsw(z) : z;
        }
    }
}

// part[block][K][N] = A^T dY over the block's 64 rows: bf16 operands
// staged in LDS, FP32 accumulation.  k_sum_parts adds the blocks' partials
// in block order, so the gradient does not depend on scheduling.  K, N
// multiples of 16 and <= 64.
template <int N>
__global__ void __launch_bounds__(256) k_gemm_wgrad(int M, int K, const float *__restrict__ A, int lda,
                                                    const float *__restrict__ dY, int ldy, float *__restrict__ part) {
    constexpr int RB = 64;
    __shared__ __align__(16) __bf16 at[64 * RB];   // A^T chunk [k][row]
    __shared__ __align__(16) __bf16 dt[N * RB];    // dY^T chunk [n][row]
    const int rb0 = blockIdx.x * RB;
    for (int i = threadIdx.x; i < RB * K; i += blockDim.x) {
        const int r = i / K, k = i % K, row = rb0 + r;
        at[k * RB + r] = (__bf16)(row < M ? A[(size_t)row * lda + k] : 0.f);
    }
    for (int i = threadIdx.x; i < RB * N; i += blockDim.x) {
        const int r = i / N, c = i % N, row = rb0 + r;
        dt[c * RB + r] = (__bf16)(row < M ? dY[(size_t)row * ldy + c] : 0.f);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int KT = K / 16;
    constexpr int NT = N / 16;
    for (int tile = wave; tile < KT * NT; tile += 4) {
        const int tk = tile / NT, tn = tile % NT;
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int r0 = 0; r0 < RB; r0 += 32) {
            const int rq = r0 + (lane >> 4) * 8;
            const bf16x8 a = *(const bf16x8 *)(at + (tk * 16 + (lane & 15)) * RB + rq);
            const bf16x8 b = *(const bf16x8 *)(dt + (tn * 16 + (lane & 15)) * RB + rq);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        const int col = tn * 16 + (lane & 15);
        float *pb = part + (size_t)blockIdx.x * K * N;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = tk * 16 + (lane >> 4) * 4 + i;
            pb[(size_t)k * N + col] = acc[i];
        }
    }
}

// out[e] += sum over blocks b (in order) of part[b][e], e < m.
__global__ void k_sum_parts(int nblocks, int m, const float *__restrict__ part, float *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; b++) s += part[(size_t)b * m + e];
    out[e] += s;
}

// part[block][c] = sum of dY[r][c] over the rows r = block, block + grid,
// ... (one thread per column; k_sum_parts finishes the sum).
__global__ void k_colsum(int M, int N, const float *__restrict__ dY, int ldy, float *__restrict__ part) {
    const int c = threadIdx.x;
    if (c >= N) return;
    float s = 0.f;
    for (int r = blockIdx.x; r < M; r += gridDim.x) s += dY[(size_t)r * ldy + c];
    part[(size_t)blockIdx.x * N + c] = s;
}

// dZ = dY * hs'(Z)
__global__ void k_act_bwd(size_t m, const float *__restrict__ dY, const float *__restrict__ Z, float *__restrict__ dZ) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < m) dZ[i] = dY[i] * hsw_grad(Z[i]);
}

// Materialise the edges of the selected particles: X [E][32], Bv [E][16]
// (12 basis components, u, 3 pad), pe [E].  One thread per selected particle.
__global__ void k_ecn_edges(EcnGraph g, EcnScales s, int np, const int *__restrict__ sel, const int *__restrict__ rowp,
                            float *__restrict__ X, float *__restrict__ Bv, int *__restrict__ pe) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int i = sel[p];
    const EcnInv iv = ecn_inv(g, s);
    const PSide pi = pside(load_state(g, i), iv);
    const int e0 = rowp[p], dg = rowp[p + 1] - e0;
    for (int k = 0; k < dg; k++) {
        const size_t slot = (size_t)k * g.n + i;
        EdgeOut eo;
        ecn_edge(g, iv, pi, true, g.nidx[slot], g.nr[slot], eo);
        float *x = X + (size_t)(e0 + k) * Mpn::KIN;
#pragma unroll
        for (int q = 0; q < Mpn::NF; q++) x[q] = eo.f[q];
#pragma unroll
        for (int q = Mpn::NF; q < Mpn::KIN; q++) x[q] = 0.f;
        float *b = Bv + (size_t)(e0 + k) * 16;
#pragma unroll
        for (int q = 0; q < 12; q++) b[q] = eo.b[q];
        b[12] = eo.u;
        b[13] = b[14] = b[15] = 0.f;
        pe[e0 + k] = p;
    }
}

// y[p] = (F[sel[p]] - g) / acc for the selected particles (F SoA [3][n]).
__global__ void k_ecn_target(int n, int np, const int *__restrict__ sel, const double *__restrict__ F, double g0,
                             double g1, double g2, float acc, float *__restrict__ y) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int i = sel[p];
    y[3 * p] = (float)((F[i] - g0) / acc);
    y[3 * p + 1] = (float)((F[(size_t)n + i] - g1) / acc);
    y[3 * p + 2] = (float)((F[2 * (size_t)n + i] - g2) / acc);
}

// pred[p] = sum over p's edges of u (o . basis), o = O[e] + b3 (O [E][4],
// the head's product without bias); with y: loss += |pred - y|^2 / np and
// D[p] = 2 (pred - y) / np.
__global__ void k_ecn_pred(int np, const int *__restrict__ rowp, const float *__restrict__ O,
                           const float *__restrict__ b3, const float *__restrict__ Bv, const float *__restrict__ y,
                           float *__restrict__ pred, float *__restrict__ D, float *__restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (p < np) {
        const float c0 = b3[0], c1 = b3[1], c2 = b3[2], c3 = b3[3];
        float a[3] = {0.f, 0.f, 0.f};
        for (int e = rowp[p]; e < rowp[p + 1]; e++) {
            const float *o = O + (size_t)e * 4, *b = Bv + (size_t)e * 16;
            const float o0 = o[0] + c0, o1 = o[1] + c1, o2 = o[2] + c2, o3 = o[3] + c3;
#pragma unroll
            for (int c = 0; c < 3; c++) a[c] += b[12] * (o0 * b[c] + o1 * b[3 + c] + o2 * b[6 + c] + o3 * b[9 + c]);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) pred[3 * p + c] = a[c];
        if (y) {
            const float sc = 2.f / (float)np;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float r = a[c] - y[3 * p + c];
                l += r * r;
                D[3 * p + c] = sc * r;
            }
        }
    }
    if (loss) {
        for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(loss, l / (float)np);
    }
}

// Head forward in FP32, one thread per edge: O[e] = H2[e] W3 + X[e][0..16) L
// (W3 [64][4], L [16][4]).
__global__ void __launch_bounds__(256) k_ecn_head_fwd(int E, const float *__restrict__ H2, const float *__restrict__ X,
                                                      const float *__restrict__ W3, const float *__restrict__ L,
                                                      float *__restrict__ O) {
    __shared__ __align__(16) float w3[Mpn::HID * Mpn::NOUT];
    __shared__ __align__(16) float wl[Mpn::NF * Mpn::NOUT];
    for (int q = threadIdx.x; q < Mpn::HID * Mpn::NOUT; q += blockDim.x) w3[q] = W3[q];
    for (int q = threadIdx.x; q < Mpn::NF * Mpn::NOUT; q += blockDim.x) wl[q] = L[q];
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const float4 *h = (const float4 *)(H2 + (size_t)e * Mpn::HID);
    float o[4] = {0.f, 0.f, 0.f, 0.f};
    const float *x = X + (size_t)e * Mpn::KIN;
    for (int f = 0; f < Mpn::NF; f++) {
        const float4 wr = ((const float4 *)wl)[f];
        o[0] += x[f] * wr.x; o[1] += x[f] * wr.y; o[2] += x[f] * wr.z; o[3] += x[f] * wr.w;
    }
    for (int q = 0; q < Mpn::HID / 4; q++) {
        const float4 hv = h[q];
        const float x[4] = {hv.x, hv.y, hv.z, hv.w};
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float4 wr = ((const float4 *)w3)[4 * q + r];
            o[0] += x[r] * wr.x; o[1] += x[r] * wr.y; o[2] += x[r] * wr.z; o[3] += x[r] * wr.w;
        }
    }
    *(float4 *)(O + (size_t)e * 4) = make_float4(o[0], o[1], o[2], o[3]);
}

// Head backward, one thread per edge: dO_k = u (B_k . D[pe]); db3 and
// dW3 [64][4] += H2^T dO (block partial sums in LDS, then atomics); and
// dZ2 = (dO W3^T) hs'(Z2) [E][64] for the hidden layers.
// Per block, into part[block][HEAD_PARTS]: dW3 [64][4], db3 [4], dL [16][4]
// (k_sum_parts adds the blocks in order).
constexpr int HEAD_PARTS = Mpn::HID * Mpn::NOUT + Mpn::NOUT + Mpn::NF * Mpn::NOUT;
__global__ void __launch_bounds__(256) k_ecn_head_bwd(int E, const int *__restrict__ pe, const float *__restrict__ Bv,
                                                      const float *__restrict__ D, const float *__restrict__ H2,
                                                      const float *__restrict__ Z2, const float *__restrict__ X,
                                                      const float *__restrict__ W3, float *__restrict__ dZ2,
                                                      float *__restrict__ part) {
    __shared__ float sdo[256][4];
    __shared__ __align__(16) float w3[Mpn::HID * Mpn::NOUT];
    for (int q = threadIdx.x; q < Mpn::HID * Mpn::NOUT; q += blockDim.x) w3[q] = W3[q];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    float d[4] = {0.f, 0.f, 0.f, 0.f};
    if (e < E) {
        const float *b = Bv + (size_t)e * 16;
        const float *dm = D + 3 * (size_t)pe[e];
#pragma unroll
        for (int k = 0; k < 4; k++) d[k] = b[12] * (b[3 * k] * dm[0] + b[3 * k + 1] * dm[1] + b[3 * k + 2] * dm[2]);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) sdo[threadIdx.x][k] = d[k];
    __syncthreads();
    if (e < E) {
        const float *z = Z2 + (size_t)e * Mpn::HID;
        float *dz = dZ2 + (size_t)e * Mpn::HID;
        for (int h = 0; h < Mpn::HID; h++) {
            const float4 wr = ((const float4 *)w3)[h];
            dz[h] = (d[0] * wr.x + d[1] * wr.y + d[2] * wr.z + d[3] * wr.w) * hsw_grad(z[h]);
        }
    }
    // thread t: dW3[t / 4][t % 4] over the block's edges, in edge order
    const int hrow = threadIdx.x >> 2, kk = threadIdx.x & 3;
    const int eb = blockIdx.x * blockDim.x;
    float *pb = part + (size_t)blockIdx.x * HEAD_PARTS;
    float s = 0.f;
    for (int q = 0; q < 256 && eb + q < E; q++) s += H2[(size_t)(eb + q) * 64 + hrow] * sdo[q][kk];
    pb[hrow * 4 + kk] = s;
    // threads 0..63: dL[t / 4][t % 4]; threads 64..67: db3
    if (threadIdx.x < Mpn::NF * Mpn::NOUT) {
        float sl = 0.f;
        for (int q = 0; q < 256 && eb + q < E; q++) sl += X[(size_t)(eb + q) * Mpn::KIN + hrow] * sdo[q][kk];
        pb[Mpn::HID * Mpn::NOUT + Mpn::NOUT + hrow * 4 + kk] = sl;
    } else if (threadIdx.x < Mpn::NF * Mpn::NOUT + Mpn::NOUT) {
        const int k = threadIdx.x - Mpn::NF * Mpn::NOUT;
        float sb = 0.f;
        for (int q = 0; q < 256 && eb + q < E; q++) sb += sdo[q][k];
        pb[Mpn::HID * Mpn::NOUT + k] = sb;
    }
}

// Adam (Kingma & Ba) on the FP32 master weights.
__global__ void k_adam(int m, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ mo,
                       float *__restrict__ ve, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const float gi = g[i];
    const float mi = b1 * mo[i] + (1.f - b1) * gi, vi = b2 * ve[i] + (1.f - b2) * gi * gi;
    mo[i] = mi;
    ve[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of the weights: W1t [64][32], W2t [64][64] (forward
// products, out-major) and W2 [64][64] (backward dH1 = dZ2 W2^T, in-major).
// The head stays FP32.
__global__ void k_ecn_bf16(const float *__restrict__ w, __bf16 *__restrict__ wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int H = Mpn::HID, K = Mpn::KIN;
    if (i < H * K) {   // W1t[o][k] = W1[k][o]
        const int o = i / K, k = i % K;
        wb[Mpn::B_W1T + i] = (__bf16)w[Mpn::OFF_W1 + k * H + o];
    }
    if (i < H * H) {
        const int o = i / H, k = i % H;
        wb[Mpn::B_W2T + i] = (__bf16)w[Mpn::OFF_W2 + k * H + o];
        wb[Mpn::B_W2 + i] = (__bf16)w[Mpn::OFF_W2 + i];
    }
}

// sum over moving particles of |F_ref - F_nn|^2 and |F_ref - g|^2 (the
// error is stated against the pair part of the acceleration)
__global__ void k_ecn_err(int n, const double *__restrict__ Fref, const double *__restrict__ Fnn,
                          const int *__restrict__ solid, double g0, double g1, double g2, double *__restrict__ acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, r = 0.0, c1 = 0.0;
    if (i < n && solid[i] == 0) {
        const double gg[3] = {g0, g1, g2};
        for (int c = 0; c < 3; c++) {
            const double a = Fref[(size_t)c * n + i], b = Fnn[(size_t)c * n + i];
            e += (a - b) * (a - b);
            r += (a - gg[c]) * (a - gg[c]);
        }
        c1 = 1.0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        e += __shfl_xor(e, o, 64);
        r += __shfl_xor(r, o, 64);
        c1 += __shfl_xor(c1, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&acc[0], e);
        atomicAdd(&acc[1], r);
        atomicAdd(&acc[2], c1);
    }
}

// ---------------------------------------------------------------------------
// Launchers.
static unsigned blocks(size_t m, int t) { return (unsigned)((m + t - 1) / t); }

void ecn_infer(hipStream_t s, const EcnGraph &g, const EcnScales &sc, const __bf16 *wb, const float *w,
               const double *gr, double *F, double cap, unsigned long long *n_capped) {
    const int nb = (g.n + 4 * ECN_PER_WAVE - 1) / (4 * ECN_PER_WAVE);
    hipLaunchKernelGGL(k_ecn_infer, xcd_grid(nb), dim3(256), 0, s, g, sc, wb, w, gr[0], gr[1], gr[2], F, cap,
                       n_capped);
}
void ecn_gemm(hipStream_t s, int M, int K, const float *A, int lda, const __bf16 *Wt, int N, const float *bias, int act,
              float *Z, float *Y, int ldy) {
    const unsigned g = blocks(M, 64);
    switch (N) {
    case 16: hipLaunchKernelGGL(k_gemm<16>, dim3(g), dim3(256), 0, s, M, K, A, lda, Wt, bias, act, Z, Y, ldy); break;
    case 32: hipLaunchKernelGGL(k_gemm<32>, dim3(g), dim3(256), 0, s, M, K, A, lda, Wt, bias, act, Z, Y, ldy); break;
    default: hipLaunchKernelGGL(k_gemm<64>, dim3(g), dim3(256), 0, s, M, K, A, lda, Wt, bias, act, Z, Y, ldy); break;
    }
}
size_t ecn_wgrad_parts(int M, int K, int N) { return (size_t)blocks(M, 64) * K * N; }
void ecn_wgrad(hipStream_t s, int M, int K, const float *A, int lda, const float *dY, int ldy, int N, float *dW,
               float *part) {
    const unsigned g = blocks(M, 64);
    switch (N) {
    case 16: hipLaunchKernelGGL(k_gemm_wgrad<16>, dim3(g), dim3(256), 0, s, M, K, A, lda, dY, ldy, part); break;
    case 32: hipLaunchKernelGGL(k_gemm_wgrad<32>, dim3(g), dim3(256), 0, s, M, K, A, lda, dY, ldy, part); break;
    default: hipLaunchKernelGGL(k_gemm_wgrad<64>, dim3(g), dim3(256), 0, s, M, K, A, lda, dY, ldy, part); break;
    }
    hipLaunchKernelGGL(k_sum_parts, dim3(blocks((size_t)K * N, 256)), dim3(256), 0, s, (int)g, K * N, part, dW);
}
void ecn_colsum(hipStream_t s, int M, int N, const float *dY, int ldy, float *db, float *part) {
    hipLaunchKernelGGL(k_colsum, dim3(256), dim3(64), 0, s, M, N, dY, ldy, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, s, 256, N, part, db);
}
void ecn_act_bwd(hipStream_t s, size_t m, const float *dY, const float *Z, float *dZ) {
    hipLaunchKernelGGL(k_act_bwd, dim3(blocks(m, 256)), dim3(256), 0, s, m, dY, Z, dZ);
}
void ecn_edges(hipStream_t s, const EcnGraph &g, const EcnScales &sc, int np, const int *sel, const int *rowp,
               float *X, float *Bv, int *pe) {
    hipLaunchKernelGGL(k_ecn_edges, dim3(blocks(np, 128)), dim3(128), 0, s, g, sc, np, sel, rowp, X, Bv, pe);
}
void ecn_target(hipStream_t s, int n, int np, const int *sel, const double *F, const double *g, float acc, float *y) {
    hipLaunchKernelGGL(k_ecn_target, dim3(blocks(np, 256)), dim3(256), 0, s, n, np, sel, F, g[0], g[1], g[2], acc, y);
}
void ecn_pred(hipStream_t s, int np, const int *rowp, const float *O, const float *b3, const float *Bv, const float *y,
              float *pred, float *D, float *loss) {
    hipLaunchKernelGGL(k_ecn_pred, dim3(blocks(np, 256)), dim3(256), 0, s, np, rowp, O, b3, Bv, y, pred, D, loss);
}
void ecn_head_fwd(hipStream_t s, int E, const float *H2, const float *X, const float *W3, const float *L, float *O) {
    hipLaunchKernelGGL(k_ecn_head_fwd, dim3(blocks(E, 256)), dim3(256), 0, s, E, H2, X, W3, L, O);
}
size_t ecn_head_parts(int E) { return (size_t)blocks(E, 256) * HEAD_PARTS; }
void ecn_head_bwd(hipStream_t s, int E, const int *pe, const float *Bv, const float *D, const float *H2,
                  const float *Z2, const float *X, const float *W3, float *dZ2, float *g_head, float *part) {
    const unsigned nb = blocks(E, 256);
    hipLaunchKernelGGL(k_ecn_head_bwd, dim3(nb), dim3(256), 0, s, E, pe, Bv, D, H2, Z2, X, W3, dZ2, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(blocks(HEAD_PARTS, 256)), dim3(256), 0, s, (int)nb, HEAD_PARTS, part, g_head);
}
void ecn_adam(hipStream_t s, int m, float *w, const float *g, float *mo, float *ve, float lr, int t) {
    const float b1 = 0.9f, b2 = 0.999f;
    hipLaunchKernelGGL(k_adam, dim3(blocks(m, 256)), dim3(256), 0, s, m, w, g, mo, ve, lr, b1, b2, 1e-8f,
                       1.f - powf(b1, (float)t), 1.f - powf(b2, (float)t));
}
void ecn_bf16(hipStream_t s, const float *w, __bf16 *wb) {
    hipLaunchKernelGGL(k_ecn_bf16, dim3(blocks(Mpn::HID * Mpn::HID, 256)), dim3(256), 0, s, w, wb);
}
void ecn_err(hipStream_t s, int n, const double *Fref, const double *Fnn, const int *solid, const double *g,
             double *acc) {
    hipLaunchKernelGGL(k_ecn_err, dim3(blocks(n, 256)), dim3(256), 0, s, n, Fref, Fnn, solid, g[0], g[1], g[2], acc);
}

}  // namespace ssa
