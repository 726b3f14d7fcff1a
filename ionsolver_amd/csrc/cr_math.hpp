// Correctly rounded float square root and division for operands in normal
// range, for the exact-arithmetic direct sums (static fields, EXACT field
// update).
//
// The correctly rounded square root and division the compiler emits
// (-fhip-fp32-correctly-rounded-divide-sqrt) are, for operands in normal
// range, v_sqrt_f32 followed by a one-ulp correction either way, and
// v_rcp_f32 followed by a Newton step and two residual corrections, wrapped
// in range scaling (v_div_scale / v_div_fmas) and special-value handling
// (v_div_fixup, class tests). Where the operands keep that wrapping the
// identity, the core sequences alone give the same bits; the callers state
// their operand ranges, and ion_psi_math_check compares both operators with
// the general ones on the device.
#pragma once
#include <hip/hip_runtime.h>

namespace ion {

// correctly rounded sqrt(x) for positive normal x of at least 2^-64
__device__ __forceinline__ float sqrt_cr(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float s_dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float s_up = __uint_as_float(__float_as_uint(s) + 1u);
    const float r_dn = __builtin_fmaf(-s_dn, s, x);
    const float r_up = __builtin_fmaf(-s_up, s, x);
    const float t = r_dn <= 0.0f ? s_dn : s;
    return r_up > 0.0f ? s_up : t;
}

// correctly rounded a / b for b in [2^-3, 2^36] and a == 0 or
// |a| in [2^-63, 2^54] (a == 0 may give a zero of the other sign)
__device__ __forceinline__ float div_cr(float a, float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    const float y1 = __builtin_fmaf(e, y0, y0);
    const float q0 = a * y1;
    const float r0 = __builtin_fmaf(-b, q0, a);
    const float q1 = __builtin_fmaf(r0, y1, q0);
    const float r1 = __builtin_fmaf(-b, q1, a);
    return __builtin_fmaf(r1, y1, q1);
}

}  // namespace ion
