// Kinetic-law programs: a C expression of a model's custom law (Reaction::setCustomLaw,
// Reaction.cpp:108-111), compiled on the host into typed stack code and evaluated per cell by
// the SSA and the deterministic integrators.
//
// The reference pastes each law's formula text into its generated kernel source
// (`_h[i+1] = _h[0] + <law>;`, Solver.cpp:843-857; `_deriv[i] = 0. + n. * <law> ...`,
// Solver.cpp:918-955; the alpha-QSS production/destruction sums, Solver.cpp:960-1010) and lets
// the OpenCL compiler type it.  Here the same text is parsed with C's rules (int, float and
// double literals; usual arithmetic conversions; left-associative operators) and every
// operation is emitted with its C type, so the program's result is the value the compiled
// expression has.  Values live on the stack as doubles, which hold every int32, float and
// double exactly; each operation rounds to its own type.
#pragma once

#include <stdint.h>

#include "gpgmp_numeric.h"

namespace gpgmp {

enum LawType : uint8_t { LT_I32 = 0, LT_F32 = 1, LT_F64 = 2 };

// One 32-bit word per instruction: op in the low 8 bits, argument above; LO_CONST is
// followed by two words (the double's low and high halves).
enum LawOp : uint32_t {
    LO_END = 0,
    LO_SPEC,     // push species arg (int32 counts in stochastic programs, float in deterministic ones)
    LO_CONST,    // push the double in the next two words
    LO_GRIDX,    // push the cell's x (int)
    LO_GRIDY,    // push the cell's y (int)
    LO_H0,       // push the running total hazard _h[0] (float)
    LO_MASK,     // push reaction arg's mask value (float 0 or 1) in this cell
    LO_I2F,      // top: (float)(int)
    LO_ADD_I, LO_SUB_I, LO_MUL_I, LO_DIV_I,
    LO_ADD_F, LO_SUB_F, LO_MUL_F, LO_DIV_F,
    LO_ADD_D, LO_SUB_D, LO_MUL_D, LO_DIV_D,
    LO_NEG_I, LO_NEG_F, LO_NEG_D,
    LO_FN_F,     // top = f(top) in float, arg: LawFn
    LO_FN_D,     // top = f(top) in double
    LO_POW_F,    // pow in float of the two top values
    LO_POW_D,
    LO_ABS_I,
    // drift/diffusivity methods (LawExtra inputs)
    LO_D2F,      // top: (float)(double)
    LO_LT_I, LO_GT_I, LO_LE_I, LO_GE_I, LO_EQ_I, LO_NE_I,   // comparisons (int result)
    LO_LT_F, LO_GT_F, LO_LE_F, LO_GE_F, LO_EQ_F, LO_NE_F,
    LO_LT_D, LO_GT_D, LO_LE_D, LO_GE_D, LO_EQ_D, LO_NE_D,
    LO_SEL,      // c ? a : b of the three top values (c deepest)
    LO_STATE,    // push species arg's count in this cell (int)
    LO_FIELD,    // push field parameter arg's value in this cell (float)
    LO_PHYSX,    // push PhysicalX / PhysicalY of this cell (float)
    LO_PHYSY,
    LO_SIMTIME,  // push PhysicalSimTime (float)
    LO_CDX,      // push getCentralDifferenceX / Y of species arg (float)
    LO_CDY,
    LO_CUR,      // push the current value of field (arg & 3) of species (arg >> 2) (float)
    LO_COUNT
};

enum LawFn : uint32_t { LF_EXP = 0, LF_LOG, LF_LOG10, LF_SQRT, LF_FABS, LF_FLOOR, LF_CEIL };

constexpr int kLawStack = 8;       // deepest stack a program may use
constexpr int kLawMaxWords = 4096;  // words of all programs of a model

// Transcendental functions of the laws, in double from the library's own exp_d/log_d (the
// same bits on the host and the GPU); float versions round the double result once.  The
// OpenCL built-ins the reference calls are implementation-defined, so these define the
// numeric contract.
GPGMP_HD double law_fn_d(uint32_t fn, double x) {
    switch (fn) {
        case LF_EXP: return exp_d(x);
        case LF_LOG: return log_d(x);
        case LF_LOG10: return log_d(x) * 0.43429448190325176;  // 1/ln(10)
        case LF_SQRT: return __builtin_sqrt(x);
        case LF_FABS: return __builtin_fabs(x);
        case LF_FLOOR: return __builtin_floor(x);
        default: return __builtin_ceil(x);
    }
}
GPGMP_HD float law_fn_f(uint32_t fn, float x) {
    switch (fn) {
        case LF_SQRT: return __builtin_sqrtf(x);
        case LF_FABS: return __builtin_fabsf(x);
        case LF_FLOOR: return __builtin_floorf(x);
        case LF_CEIL: return __builtin_ceilf(x);
        default: return (float)law_fn_d(fn, (double)x);
    }
}
// pow(x, y): an integral exponent |y| <= 64 by binary powering in double (exact for small
// integral bases: pow(2, 3) is 8); otherwise exp(y log x) for x > 0, and x = 0 and x < 0
// with an integral y by their limits.
GPGMP_HD double law_pow_d(double x, double y) {
    if (y == 0.0) return 1.0;
    if (x != x || y != y) return x + y;
    if (__builtin_floor(y) == y && __builtin_fabs(y) <= 64.0) {
        int n = (int)__builtin_fabs(y);
        double r = 1.0, b = x;
        while (n) {
            if (n & 1) r = r * b;
            b = b * b;
            n >>= 1;
        }
        return (y < 0.0) ? 1.0 / r : r;
    }
    if (x == 0.0) return (y > 0.0) ? 0.0 : __builtin_inf();
    if (x < 0.0) {
        if (__builtin_floor(y) != y) return __builtin_nan("");
        const double m = exp_d(y * log_d(-x));
        const double half = y * 0.5;
        return (__builtin_floor(half) == half) ? m : -m;
    }
    return exp_d(y * log_d(x));
}
GPGMP_HD float law_pow_f(float x, float y) { return (float)law_pow_d((double)x, (double)y); }

// The program's inputs besides the species: the cell, the running _h[0], the mask bits.
struct LawCell {
    int32_t gx, gy;
    uint32_t act;
    float h0;
};

// Inputs of a drift/diffusivity method (the generated computeDiffusionConstants kernel's
// scope, Solver.cpp:659-777): counts [S][A], field parameters [nf][A], the current
// diffusivity/drift fields [4][S][A] (X, Y diffusivity, X, Y drift), the grid and the clock.
struct LawExtra {
    const int32_t *state;
    const float *fields;
    const float *cur;
    int32_t W, H, S;
    float px, py, sim_time;
};

// Runs one program.  The opcodes are the same for every lane (the program is per model), so
// the dispatch is a uniform branch; the stack index is uniform too and selects registers
// (no private-memory array).  Returns the result as a double (exact for its type).
template <typename X, int S>
GPGMP_HD double law_run(const uint32_t *code, const X (&x)[S], const LawCell &cell, const LawExtra *ex = nullptr) {
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, top = 0;
    int sp = 0;  // values on the stack; top is the top one when sp > 0
    auto below = [&](int i) -> double {  // stack slot i (0-based, below top)
        return (i == 0) ? s0 : (i == 1) ? s1 : (i == 2) ? s2 : (i == 3) ? s3 : (i == 4) ? s4 : (i == 5) ? s5 : s6;
    };
    auto push = [&](double v) {
        if (sp > 0) {
            const int i = sp - 1;
            s0 = (i == 0) ? top : s0;
            s1 = (i == 1) ? top : s1;
            s2 = (i == 2) ? top : s2;
            s3 = (i == 3) ? top : s3;
            s4 = (i == 4) ? top : s4;
            s5 = (i == 5) ? top : s5;
            s6 = (i == 6) ? top : s6;
        }
        top = v;
        ++sp;
    };
    for (int pc = 0;; ++pc) {
#if defined(__HIP_DEVICE_COMPILE__)
        const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)code[pc]);
#else
        const uint32_t w = code[pc];
#endif
        const uint32_t op = w & 0xffu, arg = w >> 8;
        if (op == LO_END) break;
        switch (op) {
            case LO_SPEC: {
                X v = x[0];
                for (int k = 1; k < S; ++k) v = ((uint32_t)k == arg) ? x[k] : v;
                push((double)v);
                break;
            }
            case LO_CONST: {
                const uint64_t lo = code[pc + 1], hi = code[pc + 2];
                pc += 2;
                push(__builtin_bit_cast(double, lo | (hi << 32)));
                break;
            }
            case LO_GRIDX: push((double)cell.gx); break;
            case LO_GRIDY: push((double)cell.gy); break;
            case LO_H0: push((double)cell.h0); break;
            case LO_MASK: push(((cell.act >> arg) & 1u) ? 1.0 : 0.0); break;
            case LO_I2F: top = (double)(float)(int32_t)top; break;
            case LO_NEG_I: top = (double)(int32_t)(0u - (uint32_t)(int32_t)top); break;
            case LO_NEG_F: top = (double)(-(float)top); break;
            case LO_NEG_D: top = -top; break;
            case LO_FN_F: top = (double)law_fn_f(arg, (float)top); break;
            case LO_FN_D: top = law_fn_d(arg, top); break;
            case LO_ABS_I: {
                const int32_t v = (int32_t)top;
                top = (double)(int32_t)((v < 0) ? 0u - (uint32_t)v : (uint32_t)v);
                break;
            }
            case LO_D2F: top = (double)(float)top; break;
            case LO_STATE: push((double)ex->state[(size_t)arg * ex->W * ex->H + (size_t)cell.gy * ex->W + cell.gx]); break;
            case LO_FIELD: push((double)ex->fields[(size_t)arg * ex->W * ex->H + (size_t)cell.gy * ex->W + cell.gx]); break;
            case LO_PHYSX: push((double)ex->px); break;
            case LO_PHYSY: push((double)ex->py); break;
            case LO_SIMTIME: push((double)ex->sim_time); break;
            case LO_CUR: {
                const size_t A = (size_t)ex->W * ex->H;
                push((double)ex->cur[((size_t)(arg & 3u) * ex->S + (arg >> 2)) * A + (size_t)cell.gy * ex->W + cell.gx]);
                break;
            }
            case LO_CDX:
            case LO_CDY: {
                // getCentralDifferenceX/Y (cl_include/helpers.cl:39-87): the mirrored neighbour
                // at an edge; the Y form bounds dy by the grid width, as the reference does
                const int32_t *st = ex->state + (size_t)arg * ex->W * ex->H;
                const int dx = cell.gx, dy = cell.gy, W = ex->W;
                int32_t lo, hi;
                if (op == LO_CDX) {
                    lo = (dx > 0) ? st[(dx - 1) + dy * W] : st[(dx + 1) + dy * W];
                    hi = (dx < W - 1) ? st[(dx + 1) + dy * W] : st[(dx - 1) + dy * W];
                } else {
                    lo = (dy > 0) ? st[dx + (dy - 1) * W] : st[dx + (dy + 1) * W];
                    hi = (dy < W - 1) ? st[dx + (dy + 1) * W] : st[dx + (dy - 1) * W];
                }
                push((double)(float)(int32_t)((uint32_t)hi - (uint32_t)lo));
                break;
            }
            case LO_SEL: {  // c (sp-3) ? a (sp-2) : b (top)
                const double c = below(sp - 3), a = below(sp - 2);
                sp -= 2;
                top = (c != 0.0) ? a : top;
                break;
            }
            default: {  // binary: a (below) op b (top)
                const double a = below(sp - 2), b = top;
                double r;
                switch (op) {
                    case LO_ADD_I: r = (double)(int32_t)((uint32_t)(int32_t)a + (uint32_t)(int32_t)b); break;
                    case LO_SUB_I: r = (double)(int32_t)((uint32_t)(int32_t)a - (uint32_t)(int32_t)b); break;
                    case LO_MUL_I: r = (double)(int32_t)((uint32_t)(int32_t)a * (uint32_t)(int32_t)b); break;
                    case LO_DIV_I: {
                        // C truncating division; a zero divisor (undefined in C) gives 0 and
                        // INT_MIN / -1 gives INT_MIN, so no input traps
                        const int32_t ia = (int32_t)a, ib = (int32_t)b;
                        r = (ib == 0) ? 0.0 : (ib == -1) ? (double)(int32_t)(0u - (uint32_t)ia) : (double)(ia / ib);
                        break;
                    }
                    case LO_ADD_F: r = (double)((float)a + (float)b); break;
                    case LO_SUB_F: r = (double)((float)a - (float)b); break;
                    case LO_MUL_F: r = (double)((float)a * (float)b); break;
                    case LO_DIV_F: r = (double)((float)a / (float)b); break;
                    case LO_ADD_D: r = a + b; break;
                    case LO_SUB_D: r = a - b; break;
                    case LO_MUL_D: r = a * b; break;
                    case LO_DIV_D: r = a / b; break;
                    case LO_POW_F: r = (double)law_pow_f((float)a, (float)b); break;
                    case LO_POW_D: r = law_pow_d(a, b); break;
                    default: {  // comparisons: int, float and double operands are exact doubles
                        const uint32_t k = (op - LO_LT_I) % 6u;
                        const bool v = (k == 0) ? a < b : (k == 1) ? a > b : (k == 2) ? a <= b : (k == 3) ? a >= b
                                     : (k == 4) ? a == b : a != b;
                        r = v ? 1.0 : 0.0;
                        break;
                    }
                }
                --sp;
                top = r;
                break;
            }
        }
    }
    return top;
}

// Value of an expression of type t assigned to a float (`_h[i+1] = ...`, `Real _deriv = ...`).
GPGMP_HD float law_to_float(double v, uint32_t t) {
    return (t == LT_I32) ? (float)(int32_t)v : (float)v;
}

#if defined(__HIPCC__)
// One out-of-line copy of the interpreter per translation unit (the many SSA kernel variants
// call it instead of each inlining it), the species as scalar arguments so they stay in
// registers: a program's assignment to a float.
__device__ __attribute__((noinline)) inline float law_call_i(const uint32_t *code, uint32_t type, int32_t a0,
                                                             int32_t a1, int32_t a2, int32_t a3, int32_t a4,
                                                             int32_t a5, int32_t a6, int32_t a7, LawCell cell) {
    const int32_t x[8] = {a0, a1, a2, a3, a4, a5, a6, a7};
    return law_to_float(law_run<int32_t, 8>(code, x, cell), type);
}
__device__ __attribute__((noinline)) inline float law_call_f(const uint32_t *code, uint32_t type, float a0, float a1,
                                                             float a2, float a3, float a4, float a5, float a6,
                                                             float a7, LawCell cell) {
    const float x[8] = {a0, a1, a2, a3, a4, a5, a6, a7};
    return law_to_float(law_run<float, 8>(code, x, cell), type);
}
template <typename X, int S>
__device__ __forceinline__ X law_arg(const X (&x)[S], int k) {
    return (k < S) ? x[k < S ? k : 0] : (X)0;
}
template <int S>
__device__ __forceinline__ float law_call(const uint32_t *code, uint32_t type, const int32_t (&x)[S], LawCell c) {
    return law_call_i(code, type, law_arg(x, 0), law_arg(x, 1), law_arg(x, 2), law_arg(x, 3), law_arg(x, 4),
                      law_arg(x, 5), law_arg(x, 6), law_arg(x, 7), c);
}
template <int S>
__device__ __forceinline__ float law_call(const uint32_t *code, uint32_t type, const float (&x)[S], LawCell c) {
    return law_call_f(code, type, law_arg(x, 0), law_arg(x, 1), law_arg(x, 2), law_arg(x, 3), law_arg(x, 4),
                      law_arg(x, 5), law_arg(x, 6), law_arg(x, 7), c);
}
#endif

}  // namespace gpgmp
