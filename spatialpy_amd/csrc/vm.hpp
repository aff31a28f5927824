// Expression VM: custom propensities and deterministic rates of a model,
// compiled on the host from the C++ text SpatialPy generates for them
// (solver.py:160-190, 344-373; spatialpy_amd/expression.py) into a small
// stack bytecode, and interpreted here in C semantics -- int arithmetic
// wraps and truncates like the generated code's `int` expressions, double
// arithmetic is IEEE binary64 with no contraction, so the four basic
// operations, sqrt, fabs, floor, ceil, trunc, fmod, copysign, remainder and
// pow with integral exponents up to 64 give the generated function's bits;
// the other math functions are the device library's (a few ulp from the
// host libm).
//
// Instruction word: op (bits 0-5) | source (6-8) | reversed (9) | index
// (10-31).  A binary operation takes its left operand from the top of the
// stack and its right one from the source (a leaf: constant, species,
// vol, t, data function) or, for source 0, from the stack (popped); the
// reversed bit swaps the two.  Every program is checked on the host (op
// codes, operand indices, stack depth <= VM_STACK, ends in RET with one
// value) before it is uploaded.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

namespace ssa {

enum VmOp : int {
    VM_RET = 0, VM_LD = 1,
    VM_ADD_D = 2, VM_SUB_D = 3, VM_MUL_D = 4, VM_DIV_D = 5,
    VM_ADD_I = 6, VM_SUB_I = 7, VM_MUL_I = 8, VM_DIV_I = 9,
    VM_NEG_D = 10, VM_NEG_I = 11,
    VM_LT = 12, VM_LE = 13, VM_GT = 14, VM_GE = 15, VM_EQ = 16, VM_NE = 17,
    VM_AND = 18, VM_OR = 19,
    VM_CALL1 = 20, VM_CALL2 = 21,
    VM_N_OPS = 22
};
enum VmSrc : int { VM_S_STACK = 0, VM_S_CONST = 1, VM_S_X = 2, VM_S_VOL = 3, VM_S_T = 4, VM_S_DF = 5 };
constexpr int VM_STACK = 8;
constexpr int VM_FUNC1 = 26;          // ids 0..25
constexpr int VM_FUNC2_BASE = 32;     // ids 32..37
constexpr int VM_FUNC2 = 6;
constexpr int VM_MAX_STEPS = 1 << 16; // defensive bound per evaluation

// Host-side check of one program starting at `start`; returns its length
// or -1 (bad op, operand out of range, stack under/overflow, no RET).
inline int vm_check(const uint32_t *code, int n_code, int n_consts, int start, int n_x, int n_df) {
    int d = 0;
    for (int pc = start; pc < n_code; pc++) {
        const uint32_t w = code[pc];
        const int op = (int)(w & 63u), src = (int)((w >> 6) & 7u), idx = (int)(w >> 10);
        if (op >= VM_N_OPS) return -1;
        if (op == VM_RET) return d == 1 ? pc + 1 - start : -1;
        if (src > VM_S_DF) return -1;
        if (src == VM_S_CONST && idx >= n_consts) return -1;
        if (src == VM_S_X && idx >= n_x) return -1;
        if (src == VM_S_DF && idx >= n_df) return -1;
        if (op == VM_LD) {
            if (src == VM_S_STACK) return -1;
            d++;
        } else if (op == VM_NEG_D || op == VM_NEG_I || op == VM_CALL1) {
            if (d < 1 || src != VM_S_STACK) return -1;
            if (op == VM_CALL1 && idx >= VM_FUNC1) return -1;
        } else {
            if (op == VM_CALL2 && (src != VM_S_STACK || idx < VM_FUNC2_BASE || idx >= VM_FUNC2_BASE + VM_FUNC2))
                return -1;
            if (src == VM_S_STACK) {
                if (d < 2) return -1;
                d--;
            } else if (d < 1) {
                return -1;
            }
        }
        if (d > VM_STACK) return -1;
    }
    return -1;
}

__device__ __forceinline__ double vm_i(int v) { return (double)v; }
// int results wrap to 32 bits (the generated code's signed overflow on the
// reference's x86-64 build)
__device__ __forceinline__ int vm_wrap(long long v) { return (int)(unsigned)(unsigned long long)v; }

__device__ __forceinline__ double vm_call1(int f, double a) {
    switch (f) {
    case 0: return exp(a);
    case 1: return log(a);
    case 2: return log10(a);
    case 3: return log2(a);
    case 4: return log1p(a);
    case 5: return expm1(a);
    case 6: return sqrt(a);
    case 7: return sin(a);
    case 8: return cos(a);
    case 9: return tan(a);
    case 10: return asin(a);
    case 11: return acos(a);
    case 12: return atan(a);
    case 13: return sinh(a);
    case 14: return cosh(a);
    case 15: return tanh(a);
    case 16: return asinh(a);
    case 17: return acosh(a);
    case 18: return atanh(a);
    case 19: return erf(a);
    case 20: return erfc(a);
    case 21: return lgamma(a);
    case 22: return fabs(a);
    case 23: return floor(a);
    case 24: return ceil(a);
    default: return trunc(a);
    }
}
// x^n for integral |n| <= 64 rounded once from a double-double product
// (error-free products by fma, binary powering): the correctly rounded
// power, which is what the host libm's pow returns for these (its error
// stays far below half an ulp); y = 0.5 is sqrt.  Other arguments, zero,
// non-finite and subnormal results use the device pow.
__device__ __forceinline__ void vm_dd_mul(double &h, double &l, double bh, double bl) {
    const double p = h * bh;
    double e = fma(h, bh, -p);
    e += h * bl + l * bh;
    const double s = p + e;
    l = e - (s - p);
    h = s;
}
__device__ __forceinline__ double vm_pow(double x, double y) {
    if (x != 0.0 && isfinite(x) && y == trunc(y) && fabs(y) <= 64.0) {
        int n = (int)fabs(y);
        double rh = 1.0, rl = 0.0, bh = x, bl = 0.0;
        while (n) {
            if (n & 1) vm_dd_mul(rh, rl, bh, bl);
            n >>= 1;
            if (n) vm_dd_mul(bh, bl, bh, bl);
        }
        double r = rh;
        if (y < 0.0) {
            const double q = 1.0 / rh;
            const double res = fma(-q, rh, 1.0) - q * rl;   // 1 - q*(rh + rl)
            r = q + q * res;
        }
        if (isfinite(r) && fabs(r) >= 0x1p-1000 && fabs(r) <= 0x1p1000) return r;
    }
    if (y == 0.5 && x > 0.0 && isfinite(x)) return sqrt(x);
    return pow(x, y);
}
__device__ __forceinline__ double vm_call2(int f, double a, double b) {
    switch (f - VM_FUNC2_BASE) {
    case 0: return vm_pow(a, b);
    case 1: return atan2(a, b);
    case 2: return hypot(a, b);
    case 3: return fmod(a, b);
    case 4: return copysign(a, b);
    default: return remainder(a, b);
    }
}

__device__ __forceinline__ double vm_apply(int op, int idx, double l, double r) {
    switch (op) {
    case VM_ADD_D: return l + r;
    case VM_SUB_D: return l - r;
    case VM_MUL_D: return l * r;
    case VM_DIV_D: return l / r;
    case VM_ADD_I: return vm_i(vm_wrap((long long)(int)l + (long long)(int)r));
    case VM_SUB_I: return vm_i(vm_wrap((long long)(int)l - (long long)(int)r));
    case VM_MUL_I: return vm_i(vm_wrap((long long)(int)l * (long long)(int)r));
    case VM_DIV_I: {
        const int a = (int)l, b = (int)r;
        // the reference's program dies with SIGFPE on these two; 0 / INT_MIN here
        if (b == 0) return 0.0;
        if (a == INT_MIN && b == -1) return vm_i(INT_MIN);
        return vm_i(a / b);
    }
    case VM_LT: return l < r ? 1.0 : 0.0;
    case VM_LE: return l <= r ? 1.0 : 0.0;
    case VM_GT: return l > r ? 1.0 : 0.0;
    case VM_GE: return l >= r ? 1.0 : 0.0;
    case VM_EQ: return l == r ? 1.0 : 0.0;
    case VM_NE: return l != r ? 1.0 : 0.0;
    case VM_AND: return (l != 0.0 && r != 0.0) ? 1.0 : 0.0;
    case VM_OR: return (l != 0.0 || r != 0.0) ? 1.0 : 0.0;
    default: return vm_call2(idx, l, r);   // VM_CALL2
    }
}

// One evaluation.  XT = uint32_t: species counts read as `int` (a
// propensity, const int *x); XT = double: concentrations (a deterministic
// rate, const double *x).  The stack lives in registers (constant indices
// only).
template <class XT>
__device__ __forceinline__ double vm_run(const uint32_t *code, int pc, const double *K, const XT *x, double vol,
                                         double t, const double *df) {
    double s[VM_STACK];
#pragma unroll
    for (int k = 0; k < VM_STACK; k++) s[k] = 0.0;
    for (int it = 0; it < VM_MAX_STEPS; it++) {
        const uint32_t w = code[pc++];
        const int op = (int)(w & 63u), src = (int)((w >> 6) & 7u), idx = (int)(w >> 10);
        if (op == VM_RET) return s[0];
        if (op == VM_NEG_D) { s[0] = -s[0]; continue; }
        if (op == VM_NEG_I) { s[0] = vm_i(vm_wrap(-(long long)(int)s[0])); continue; }
        if (op == VM_CALL1) { s[0] = vm_call1(idx, s[0]); continue; }
        double b;
        switch (src) {
        case VM_S_CONST: b = K[idx]; break;
        case VM_S_X:
            if constexpr (sizeof(XT) == 4) b = vm_i((int)x[idx]);
            else b = (double)x[idx];
            break;
        case VM_S_VOL: b = vol; break;
        case VM_S_T: b = t; break;
        case VM_S_DF: b = df[idx]; break;
        default:
            b = s[0];
#pragma unroll
            for (int k = 0; k < VM_STACK - 1; k++) s[k] = s[k + 1];
            break;
        }
        if (op == VM_LD) {
#pragma unroll
            for (int k = VM_STACK - 1; k > 0; k--) s[k] = s[k - 1];
            s[0] = b;
            continue;
        }
        const bool rev = (w >> 9) & 1u;
        const double l = rev ? b : s[0], r = rev ? s[0] : b;
        s[0] = vm_apply(op, idx, l, r);
    }
    return __builtin_nan("");
}

// Out-of-line entry points (one copy per kernel that calls them keeps the
// callers' register allocation untouched).
static __device__ __noinline__ double vm_eval_u32(const uint32_t *code, int pc, const double *K, const uint32_t *x,
                                                  double vol, double t, const double *df) {
    return vm_run<uint32_t>(code, pc, K, x, vol, t, df);
}
static __device__ __noinline__ double vm_eval_f64(const uint32_t *code, int pc, const double *K, const double *x,
                                                  double vol, double t, const double *df) {
    return vm_run<double>(code, pc, K, x, vol, t, df);
}

}  // namespace ssa
