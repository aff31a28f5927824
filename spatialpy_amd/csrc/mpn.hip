// Message-passing surrogate for the SDPD accelerations on gfx950.
//
// Network (hidden width 64, three message-passing layers):
//   encoder   H0 = silu(X0 We + be)                     X0: 10 node features (padded to 32)
//   layer l   AB = H W1ab                               [n][128]: A_i = H_i W1a, B_j = H_j W1b
//             S_i = sum_j silu(A_i + B_j + e_ij W1c + b1)    e_ij = (dx/h, r/h)
//             M = S W2 + deg * b2
//             U = silu([H | M] Wn1 + bn1)
//             H' = H + U Wn2 + bn2
//   decoder   out = H3 Wd + bd                          3 accelerations (normalised)
// This is the edge MLP 132->64->64 (inputs [H_i, H_j, e_ij]) with a sum
// aggregation and the node MLP 128->64->64: the edge MLP's first layer is
// split into per-node products (A, B) and, its second layer being linear,
// the sum over neighbours is taken before it, so every dense product is a
// node-level GEMM on the matrix cores.
//
// GEMMs: v_mfma_f32_16x16x32_bf16, bf16 operands (weights stored in bf16,
// activations rounded to bf16 as they are loaded), FP32 accumulation.  One
// wavefront computes 16 rows x all output columns; the weight matrix sits in
// LDS transposed, so a lane's 8 k-values are one 16-byte read.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "engine.hpp"

namespace ssa {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s + z * s * (1.0f - s);
}

// ---------------------------------------------------------------------------
// Y[M][N] = act(A[M][K] W[K][N] + bias * rowscale) (+ R), A from two sources
// split at column K1 (A0: rows of stride lda0 holding columns [0, K1), A1:
// stride lda1 holding [K1, K)); Wt: bf16 [N][K] (the
// weights transposed).  Z (optional) receives the pre-activation.  K % 32 == 0,
// N % 16 == 0, N <= 128, K <= 128.
template <int N>
__global__ void __launch_bounds__(256) k_gemm(int M, int K, int K1, const float *__restrict__ A0, int lda0,
                                              const float *__restrict__ A1, int lda1, const __bf16 *__restrict__ Wt,
                                              const float *__restrict__ bias, const int *__restrict__ rowscale,
                                              int act, const float *__restrict__ R, float *__restrict__ Z,
                                              float *__restrict__ Y) {
    __shared__ __align__(16) __bf16 w[N * 128];
    for (int i = threadIdx.x; i < N * K; i += blockDim.x) w[i] = Wt[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * 4 + wave) * 16;
    if (r0 >= M) return;
    const int row = r0 + (lane & 15);
    const int kq = (lane >> 4) * 8;          // this lane's 8 k-values within a 32-wide step
    constexpr int NT = N / 16;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        const int kc = k0 + kq;
        bf16x8 a;
        if (row < M) {
            const float *src = kc < K1 ? A0 + (size_t)row * lda0 + kc : A1 + (size_t)row * lda1 + (kc - K1);
            const float4 lo = *(const float4 *)src, hi = *(const float4 *)(src + 4);
            a = bf16x8{(__bf16)lo.x, (__bf16)lo.y, (__bf16)lo.z, (__bf16)lo.w,
                       (__bf16)hi.x, (__bf16)hi.y, (__bf16)hi.z, (__bf16)hi.w};
        } else {
            a = bf16x8{};
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int col = t * 16 + (lane & 15);
            const bf16x8 b = *(const bf16x8 *)(w + col * K + kc);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
    }
    // D layout: lane holds rows 4*(lane/16)+i, column lane%16 of each tile
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int col = t * 16 + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = r0 + (lane >> 4) * 4 + i;
            if (r >= M) continue;
            float z = acc[t][i];
            if (bias) z += bias[col] * (rowscale ? (float)rowscale[r] : 1.0f);
            if (Z) Z[(size_t)r * N + col] = z;
            float y = act ? silu(z) : z;
            if (R) y += R[(size_t)r * N + col];
            Y[(size_t)r * N + col] = y;
        }
    }
}

// dW[K][N] += A^T dY over the rows (A split at K1 as above), bf16 operands
// staged in LDS, FP32 accumulation; one block per 64 rows, partial tiles
// added with atomics.  K, N multiples of 16 and <= 128.
template <int N>
__global__ void __launch_bounds__(256) k_gemm_wgrad(int M, int K, int K1, const float *__restrict__ A0, int lda0,
                                                    const float *__restrict__ A1, int lda1,
                                                    const float *__restrict__ dY, int ldy, float *__restrict__ dW) {
    constexpr int RB = 64;
    __shared__ __align__(16) __bf16 at[128 * RB];   // A^T chunk [k][row]
    __shared__ __align__(16) __bf16 dt[N * RB];     // dY^T chunk [n][row]
    const int rb0 = blockIdx.x * RB;
    for (int i = threadIdx.x; i < RB * K; i += blockDim.x) {
        const int r = i / K, k = i % K, row = rb0 + r;
        float v = 0.f;
        if (row < M) v = k < K1 ? A0[(size_t)row * lda0 + k] : A1[(size_t)row * lda1 + (k - K1)];
        at[k * RB + r] = (__bf16)v;
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
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = tk * 16 + (lane >> 4) * 4 + i;
            atomicAdd(&dW[(size_t)k * N + col], acc[i]);
        }
    }
}

// db[c] += sum_r w_r dY[r][c] (w_r = rowscale or 1), one thread per column.
__global__ void k_colsum(int M, int N, const float *__restrict__ dY, int ldy, const int *__restrict__ rowscale,
                         float *__restrict__ db) {
    const int c = threadIdx.x;
    if (c >= N) return;
    float s = 0.f;
    for (int r = blockIdx.x; r < M; r += gridDim.x) s += dY[(size_t)r * ldy + c] * (rowscale ? (float)rowscale[r] : 1.f);
    atomicAdd(&db[c], s);
}

// ---------------------------------------------------------------------------
// Edge features of edge (i, j): dx/h and r/h from the current positions.
struct EdgeIn {
    const float *x;      // [n][3] positions
    float inv_h;
    int n, W;
    const int *deg, *nidx;   // ELL [W][n]
};
__device__ __forceinline__ void edge_feat(const EdgeIn &g, int i, int j, float e[4]) {
    const float dx = (g.x[3 * i] - g.x[3 * j]) * g.inv_h, dy = (g.x[3 * i + 1] - g.x[3 * j + 1]) * g.inv_h,
                dz = (g.x[3 * i + 2] - g.x[3 * j + 2]) * g.inv_h;
    e[0] = dx; e[1] = dy; e[2] = dz;
    e[3] = sqrtf(dx * dx + dy * dy + dz * dz);
}

// S_i[f] = sum_j silu(AB_i[f] + AB_j[64+f] + e_ij . W1c[:, f] + b1[f]); one
// wavefront per node, lane f.
__global__ void __launch_bounds__(256) k_edge_fwd(EdgeIn g, const float *__restrict__ AB, const float *__restrict__ W1c,
                                                  const float *__restrict__ b1, float *__restrict__ Sout) {
    const int f = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= g.n) return;
    const float a = AB[(size_t)i * 128 + f] + b1[f];
    const float w0 = W1c[f], w1 = W1c[64 + f], w2 = W1c[128 + f], w3 = W1c[192 + f];
    float s = 0.f;
    const int dg = g.deg[i];
    for (int k = 0; k < dg; k++) {
        const int j = g.nidx[(size_t)k * g.n + i];
        float e[4];
        edge_feat(g, i, j, e);
        const float z = a + AB[(size_t)j * 128 + 64 + f] + e[0] * w0 + e[1] * w1 + e[2] * w2 + e[3] * w3;
        s += silu(z);
    }
    Sout[(size_t)i * 64 + f] = s;
}

// Backward of k_edge_fwd: dAB_i[f] += g, dAB_j[64+f] += g (atomic), dW1c, db1,
// with g = dS_i[f] silu'(z_ij[f]).
__global__ void __launch_bounds__(256) k_edge_bwd(EdgeIn g, const float *__restrict__ AB, const float *__restrict__ W1c,
                                                  const float *__restrict__ b1, const float *__restrict__ dS,
                                                  float *__restrict__ dAB, float *__restrict__ dW1c,
                                                  float *__restrict__ db1) {
    const int f = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= g.n) return;
    const float a = AB[(size_t)i * 128 + f] + b1[f];
    const float w0 = W1c[f], w1 = W1c[64 + f], w2 = W1c[128 + f], w3 = W1c[192 + f];
    const float ds = dS[(size_t)i * 64 + f];
    float ga = 0.f, gw0 = 0.f, gw1 = 0.f, gw2 = 0.f, gw3 = 0.f;
    const int dg = g.deg[i];
    for (int k = 0; k < dg; k++) {
        const int j = g.nidx[(size_t)k * g.n + i];
        float e[4];
        edge_feat(g, i, j, e);
        const float z = a + AB[(size_t)j * 128 + 64 + f] + e[0] * w0 + e[1] * w1 + e[2] * w2 + e[3] * w3;
        const float gz = ds * silu_grad(z);
        ga += gz;
        atomicAdd(&dAB[(size_t)j * 128 + 64 + f], gz);
        gw0 += gz * e[0]; gw1 += gz * e[1]; gw2 += gz * e[2]; gw3 += gz * e[3];
    }
    atomicAdd(&dAB[(size_t)i * 128 + f], ga);
    atomicAdd(&dW1c[f], gw0);
    atomicAdd(&dW1c[64 + f], gw1);
    atomicAdd(&dW1c[128 + f], gw2);
    atomicAdd(&dW1c[192 + f], gw3);
    atomicAdd(&db1[f], ga);
}

// out[i][c] = H[i] . Wd[:, c] + bd[c]   (64 -> 3)
__global__ void k_decode(int n, const float *__restrict__ H, const float *__restrict__ Wd, const float *__restrict__ bd,
                         float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float o0 = bd[0], o1 = bd[1], o2 = bd[2];
    const float *h = H + (size_t)i * 64;
    for (int f = 0; f < 64; f++) {
        const float v = h[f];
        o0 += v * Wd[f * 3]; o1 += v * Wd[f * 3 + 1]; o2 += v * Wd[f * 3 + 2];
    }
    out[3 * i] = o0; out[3 * i + 1] = o1; out[3 * i + 2] = o2;
}

// loss = mean((out - y)^2) over n*3 values; dout = 2 (out - y) / (3n);
// dH = dout Wd^T; dWd += H^T dout; dbd += sum dout
__global__ void k_decode_bwd(int n, const float *__restrict__ H, const float *__restrict__ Wd,
                             const float *__restrict__ out, const float *__restrict__ y, float *__restrict__ dH,
                             float *__restrict__ dWd, float *__restrict__ dbd, float *__restrict__ loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ float red[3 * 64 + 4];
    for (int q = threadIdx.x; q < 3 * 64 + 4; q += blockDim.x) red[q] = 0.f;
    __syncthreads();
    if (i < n) {
        const float sc = 2.0f / (3.0f * (float)n);
        float d[3], l = 0.f;
        for (int c = 0; c < 3; c++) {
            const float r = out[3 * i + c] - y[3 * i + c];
            l += r * r;
            d[c] = sc * r;
        }
        atomicAdd(&red[192], l / (3.0f * (float)n));
        for (int c = 0; c < 3; c++) atomicAdd(&red[193 + c], d[c]);
        const float *h = H + (size_t)i * 64;
        for (int f = 0; f < 64; f++) {
            dH[(size_t)i * 64 + f] = d[0] * Wd[f * 3] + d[1] * Wd[f * 3 + 1] + d[2] * Wd[f * 3 + 2];
            const float v = h[f];
            atomicAdd(&red[f * 3], v * d[0]);
            atomicAdd(&red[f * 3 + 1], v * d[1]);
            atomicAdd(&red[f * 3 + 2], v * d[2]);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < 192; q += blockDim.x) atomicAdd(&dWd[q], red[q]);
    if (threadIdx.x < 3) atomicAdd(&dbd[threadIdx.x], red[193 + threadIdx.x]);
    if (threadIdx.x == 0) atomicAdd(loss, red[192]);
}

// dZ = dY * silu'(Z)
__global__ void k_silu_bwd(size_t m, const float *__restrict__ dY, const float *__restrict__ Z, float *__restrict__ dZ) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < m) dZ[i] = dY[i] * silu_grad(Z[i]);
}
// y[r][c] += x[r*ldx + c] for a [rows][cols] y
__global__ void k_add(int rows, int cols, float *__restrict__ y, const float *__restrict__ x, int ldx) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    const size_t r = i / cols, c = i % cols;
    y[i] += x[r * ldx + c];
}

// Adam (Kingma & Ba) on the FP32 master weights, then the bf16 transposed
// copies used by the GEMMs are refreshed.
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
// Wt[c][k] = bf16(W[k][c]) for a K x N matrix (rows padded to Kp with zeros)
__global__ void k_to_bf16_t(int K, int Kp, int N, const float *__restrict__ W, __bf16 *__restrict__ Wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Kp * N) return;
    const int c = i / Kp, k = i % Kp;
    Wt[i] = (__bf16)(k < K ? W[(size_t)k * N + c] : 0.f);
}
// W^T as a bf16 [K][N] matrix in the layout k_gemm wants for computing
// dX = dY W^T: the "transposed weights" of that product are W itself.
__global__ void k_to_bf16(int K, int N, const float *__restrict__ W, __bf16 *__restrict__ Wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K * N) Wb[i] = (__bf16)W[i];
}

// Node features from the engine's state: v/v_scale (3), (rho/rho0 - 1)/rho_scale,
// solid flag, nu/nu_scale, 1, zeros to 32.  Positions to FP32.
__global__ void k_mpn_features(int n, const double *__restrict__ x, const double *__restrict__ v,
                               const double *__restrict__ vt, const double *__restrict__ rho,
                               const double *__restrict__ nu, const int *__restrict__ solid, float rho0,
                               float v_scale, float rho_scale, float nu_scale, float *__restrict__ X0,
                               float *__restrict__ xf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float *o = X0 + (size_t)i * 32;
    for (int c = 0; c < 3; c++) {
        const double vc = v[(size_t)c * n + i];
        o[c] = (float)vc / v_scale;
        o[7 + c] = (float)(vt[(size_t)c * n + i] - vc) / v_scale;
    }
    o[3] = ((float)rho[i] / rho0 - 1.0f) / rho_scale;
    o[4] = solid[i] ? 1.0f : 0.0f;
    o[5] = (float)nu[i] / nu_scale;
    o[6] = 1.0f;
    for (int c = 10; c < 32; c++) o[c] = 0.f;
    xf[3 * i] = (float)x[i];
    xf[3 * i + 1] = (float)x[(size_t)n + i];
    xf[3 * i + 2] = (float)x[2 * (size_t)n + i];
}

// F = g + acc_scale * out for the moving particles (fixed ones keep gravity,
// which no later substep reads for them).
__global__ void k_mpn_apply(int n, int dim, const float *__restrict__ out, const int *__restrict__ solid, double g0,
                            double g1, double g2, float acc_scale, double *__restrict__ F) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool mv = solid[i] == 0;
    // components past the dimension keep gravity alone, as in the pair kernel
    F[i] = g0 + (mv ? (double)(acc_scale * out[3 * i]) : 0.0);
    F[(size_t)n + i] = g1 + (mv && dim > 1 ? (double)(acc_scale * out[3 * i + 1]) : 0.0);
    F[2 * (size_t)n + i] = g2 + (mv && dim > 2 ? (double)(acc_scale * out[3 * i + 2]) : 0.0);
}

// sum over moving particles of |F_ref - F_nn|^2 and |F_ref - g|^2 (the
// error is stated against the pair part of the acceleration)
__global__ void k_mpn_err(int n, const double *__restrict__ Fref, const double *__restrict__ Fnn,
                          const int *__restrict__ solid, double g0, double g1, double g2, double *__restrict__ acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, r = 0.0;
    if (i < n && solid[i] == 0) {
        const double gg[3] = {g0, g1, g2};
        for (int c = 0; c < 3; c++) {
            const double a = Fref[(size_t)c * n + i], b = Fnn[(size_t)c * n + i];
            e += (a - b) * (a - b);
            r += (a - gg[c]) * (a - gg[c]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) { e += __shfl_xor(e, o, 64); r += __shfl_xor(r, o, 64); }
    if ((threadIdx.x & 63) == 0) { atomicAdd(&acc[0], e); atomicAdd(&acc[1], r); }
}

// ---------------------------------------------------------------------------
// Host side.
static unsigned blocks(size_t m, int t) { return (unsigned)((m + t - 1) / t); }

void mpn_gemm(hipStream_t s, int M, int K, int K1, const float *A0, int lda0, const float *A1, int lda1,
              const __bf16 *Wt, int N, const float *bias, const int *rowscale, int act, const float *R, float *Z,
              float *Y) {
    const unsigned g = blocks(M, 64);
    switch (N) {
    case 16: hipLaunchKernelGGL(k_gemm<16>, dim3(g), dim3(256), 0, s, M, K, K1, A0, lda0, A1, lda1, Wt, bias, rowscale, act, R, Z, Y); break;
    case 32: hipLaunchKernelGGL(k_gemm<32>, dim3(g), dim3(256), 0, s, M, K, K1, A0, lda0, A1, lda1, Wt, bias, rowscale, act, R, Z, Y); break;
    case 64: hipLaunchKernelGGL(k_gemm<64>, dim3(g), dim3(256), 0, s, M, K, K1, A0, lda0, A1, lda1, Wt, bias, rowscale, act, R, Z, Y); break;
    default: hipLaunchKernelGGL(k_gemm<128>, dim3(g), dim3(256), 0, s, M, K, K1, A0, lda0, A1, lda1, Wt, bias, rowscale, act, R, Z, Y); break;
    }
}
void mpn_wgrad(hipStream_t s, int M, int K, int K1, const float *A0, int lda0, const float *A1, int lda1,
               const float *dY, int ldy, int N, float *dW) {
    const unsigned g = blocks(M, 64);
    switch (N) {
    case 16: hipLaunchKernelGGL(k_gemm_wgrad<16>, dim3(g), dim3(256), 0, s, M, K, K1, A0, lda0, A1, lda1, dY, ldy, dW); break;
    case 32: hipLaunchKernelGGL(k_gemm_wgrad<32>, dim3(g), dim3(256), 0, s, M, K, K1, A0, lda0, A1, lda1, dY, ldy, dW); break;
    case 64: hipLaunchKernelGGL(k_gemm_wgrad<64>, dim3(g), dim3(256), 0, s, M, K, K1, A0, lda0, A1, lda1, dY, ldy, dW); break;
    default: hipLaunchKernelGGL(k_gemm_wgrad<128>, dim3(g), dim3(256), 0, s, M, K, K1, A0, lda0, A1, lda1, dY, ldy, dW); break;
    }
}
void mpn_colsum(hipStream_t s, int M, int N, const float *dY, int ldy, const int *rowscale, float *db) {
    hipLaunchKernelGGL(k_colsum, dim3(256), dim3(128), 0, s, M, N, dY, ldy, rowscale, db);
}
void mpn_edge_fwd(hipStream_t s, const MpnGraph &g, const float *AB, const float *W1c, const float *b1, float *S) {
    EdgeIn e{g.x, g.inv_h, g.n, g.W, g.deg, g.nidx};
    hipLaunchKernelGGL(k_edge_fwd, dim3(blocks(g.n, 4)), dim3(256), 0, s, e, AB, W1c, b1, S);
}
void mpn_edge_bwd(hipStream_t s, const MpnGraph &g, const float *AB, const float *W1c, const float *b1, const float *dS,
                  float *dAB, float *dW1c, float *db1) {
    EdgeIn e{g.x, g.inv_h, g.n, g.W, g.deg, g.nidx};
    hipLaunchKernelGGL(k_edge_bwd, dim3(blocks(g.n, 4)), dim3(256), 0, s, e, AB, W1c, b1, dS, dAB, dW1c, db1);
}
void mpn_decode(hipStream_t s, int n, const float *H, const float *Wd, const float *bd, float *out) {
    hipLaunchKernelGGL(k_decode, dim3(blocks(n, 256)), dim3(256), 0, s, n, H, Wd, bd, out);
}
void mpn_decode_bwd(hipStream_t s, int n, const float *H, const float *Wd, const float *out, const float *y, float *dH,
                    float *dWd, float *dbd, float *loss) {
    hipLaunchKernelGGL(k_decode_bwd, dim3(blocks(n, 256)), dim3(256), 0, s, n, H, Wd, out, y, dH, dWd, dbd, loss);
}
void mpn_silu_bwd(hipStream_t s, size_t m, const float *dY, const float *Z, float *dZ) {
    hipLaunchKernelGGL(k_silu_bwd, dim3(blocks(m, 256)), dim3(256), 0, s, m, dY, Z, dZ);
}
void mpn_add(hipStream_t s, int rows, int cols, float *y, const float *x, int ldx) {
    hipLaunchKernelGGL(k_add, dim3(blocks((size_t)rows * cols, 256)), dim3(256), 0, s, rows, cols, y, x, ldx);
}
void mpn_adam(hipStream_t s, int m, float *w, const float *g, float *mo, float *ve, float lr, int t) {
    const float b1 = 0.9f, b2 = 0.999f;
    hipLaunchKernelGGL(k_adam, dim3(blocks(m, 256)), dim3(256), 0, s, m, w, g, mo, ve, lr, b1, b2, 1e-8f,
                       1.f - powf(b1, (float)t), 1.f - powf(b2, (float)t));
}
void mpn_to_bf16_t(hipStream_t s, int K, int Kp, int N, const float *W, __bf16 *Wt) {
    hipLaunchKernelGGL(k_to_bf16_t, dim3(blocks((size_t)Kp * N, 256)), dim3(256), 0, s, K, Kp, N, W, Wt);
}
void mpn_to_bf16(hipStream_t s, int K, int N, const float *W, __bf16 *Wb) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks((size_t)K * N, 256)), dim3(256), 0, s, K, N, W, Wb);
}
void mpn_features(hipStream_t s, int n, const double *x, const double *v, const double *vt, const double *rho,
                  const double *nu, const int *solid, float rho0, float v_scale, float rho_scale, float nu_scale,
                  float *X0, float *xf) {
    hipLaunchKernelGGL(k_mpn_features, dim3(blocks(n, 256)), dim3(256), 0, s, n, x, v, vt, rho, nu, solid, rho0,
                       v_scale, rho_scale, nu_scale, X0, xf);
}

// y[i] = (F_i - g) / acc_scale for moving particles, 0 for fixed ones
__global__ void k_mpn_target(int n, const double *__restrict__ F, const int *__restrict__ solid, double g0, double g1,
                             double g2, float acc_scale, float *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool mv = solid[i] == 0;
    y[3 * i] = mv ? (float)((F[i] - g0) / acc_scale) : 0.f;
    y[3 * i + 1] = mv ? (float)((F[(size_t)n + i] - g1) / acc_scale) : 0.f;
    y[3 * i + 2] = mv ? (float)((F[2 * (size_t)n + i] - g2) / acc_scale) : 0.f;
}
void mpn_target(hipStream_t s, int n, const double *F, const int *solid, const double *g, float acc_scale, float *y) {
    hipLaunchKernelGGL(k_mpn_target, dim3(blocks(n, 256)), dim3(256), 0, s, n, F, solid, g[0], g[1], g[2], acc_scale, y);
}
void mpn_apply(hipStream_t s, int n, int dim, const float *out, const int *solid, const double *g, float acc_scale,
               double *F) {
    hipLaunchKernelGGL(k_mpn_apply, dim3(blocks(n, 256)), dim3(256), 0, s, n, dim, out, solid, g[0], g[1], g[2],
                       acc_scale, F);
}
void mpn_err(hipStream_t s, int n, const double *Fref, const double *Fnn, const int *solid, const double *g,
             double *acc) {
    hipLaunchKernelGGL(k_mpn_err, dim3(blocks(n, 256)), dim3(256), 0, s, n, Fref, Fnn, solid, g[0], g[1], g[2], acc);
}

}  // namespace ssa
