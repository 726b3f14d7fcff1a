// Device-side lattice arithmetic for the gfx950 kernels: velocity-set tables,
// DDF storage codecs, equilibria, moments and Guo forcing. Every function is
// evaluated in the operation order of the reference kernels
// (src/kernels/sim_kernels.cl) so that results agree exactly with a strict
// IEEE restatement; the sources must be compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ion {

// ---------------------------------------------------------------------------
// Arithmetic type R: float for the product kernels, double for the FP64
// instantiation (the error reference of the storage formats). Every function
// below runs the same operation sequence in either type.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaR(double a, double b, double c) { return __builtin_fma(a, b, c); }
// "1/3" as written in the forcing terms (0.33333334f) and its FP64 value
template <class R> __device__ __forceinline__ R third();
template <> __device__ __forceinline__ float third<float>() { return 0.33333334f; }
template <> __device__ __forceinline__ double third<double>() { return 1.0 / 3.0; }
// OpenCL clamp(x, lo, hi) = fmin(fmax(x, lo), hi)
__device__ __forceinline__ float clampR(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ double clampR(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return clampR(x, lo, hi); }

// ---------------------------------------------------------------------------
// Velocity sets, reference direction ordering (sim_kernels.cl:336-339,
// neighbor order :277-286, weights src/lbm/domain.rs:757-768: 1/d0, 1/ds,
// 1/de evaluated in the arithmetic type).
// ---------------------------------------------------------------------------
template <int Q> struct VS;
template <> struct VS<9> {
    static constexpr int dims = 2;
    static constexpr int cx[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
    static constexpr int cy[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
    static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr int cls[9] = {0, 1, 1, 1, 1, 2, 2, 2, 2};
    static constexpr double d0 = 2.25, ds = 9.0, de = 36.0;
};
template <> struct VS<15> {
    static constexpr int dims = 3;
    static constexpr int cx[15] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    static constexpr int cy[15] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1};
    static constexpr int cz[15] = {0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1, 1, -1};
    static constexpr int cls[15] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2};
    static constexpr double d0 = 4.5, ds = 9.0, de = 72.0;
};
template <> struct VS<19> {
    static constexpr int dims = 3;
    static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
    static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
    static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
    static constexpr int cls[19] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
    static constexpr double d0 = 3.0, ds = 18.0, de = 36.0;
};

template <int Q, class R> __device__ __forceinline__ constexpr R weight_of_class(int c) {
    return c == 0 ? R(1) / R(VS<Q>::d0) : (c == 1 ? R(1) / R(VS<Q>::ds) : R(1) / R(VS<Q>::de));
}
template <int Q, class R> __device__ __forceinline__ constexpr R weight(int i) {
    return weight_of_class<Q, R>(VS<Q>::cls[i]);
}

// Signed sum c_i . (X, Y, Z) over the non-zero components, left to right,
// the first term carrying its own sign (e.g. "-ux+uy+uz"). Multiplications by
// the unit direction components are exact and are not performed.
template <int Q, int I, class R> __device__ __forceinline__ R proj(R X, R Y, R Z) {
    constexpr int a = VS<Q>::cx[I], b = VS<Q>::cy[I], c = VS<Q>::cz[I];
    R s = R(0);
    bool first = true;
    if (a != 0) { s = a > 0 ? X : -X; first = false; }
    if (b != 0) { s = first ? (b > 0 ? Y : -Y) : (b > 0 ? s + Y : s - Y); first = false; }
    if (c != 0) { s = first ? (c > 0 ? Z : -Z) : (c > 0 ? s + Z : s - Z); }
    return s;
}

template <int Q, int I> struct Unroll {
    template <class F> __device__ __forceinline__ static void run(F&& f) {
        Unroll<Q, I - 1>::run(f);
        f(std::integral_constant<int, I - 1>{});
    }
};
template <int Q> struct Unroll<Q, 0> {
    template <class F> __device__ __forceinline__ static void run(F&&) {}
};
// static_for over i = 0..N-1 with a compile-time index
template <int N, class F> __device__ __forceinline__ void static_for(F&& f) { Unroll<N, N>::run(f); }

// calculate_f_eq, sim_kernels.cl:155-207
template <int Q, class R> __device__ __forceinline__ void f_eq(R rho, R ux, R uy, R uz, R* feq) {
    const R c3 = R(-3) * ((ux * ux + uy * uy) + uz * uz);
    const R rhom1 = rho - R(1);
    ux *= R(3);
    uy *= R(3);
    uz *= R(3);
    feq[0] = weight_of_class<Q, R>(0) * fmaR(rho, R(0.5) * c3, rhom1);
    const R ws = weight_of_class<Q, R>(1), we = weight_of_class<Q, R>(2);
    const R rhos = ws * rho, rhoe = we * rho;
    const R rhom1s = ws * rhom1, rhom1e = we * rhom1;
    static_for<(Q - 1) / 2>([&](auto p) {
        constexpr int i = 2 * decltype(p)::value + 1;
        const R s = proj<Q, i, R>(ux, uy, uz);
        const R rw = VS<Q>::cls[i] == 1 ? rhos : rhoe;
        const R rm = VS<Q>::cls[i] == 1 ? rhom1s : rhom1e;
        const R sq = fmaR(s, s, c3);
        feq[i] = fmaR(rw, fmaR(R(0.5), sq, s), rm);
        feq[i + 1] = fmaR(rw, fmaR(R(0.5), sq, -s), rm);
    });
}

// calculate_rho_u, sim_kernels.cl:208-233: density summed in index order then
// +1; each velocity component alternates + and - over the direction pairs
// with a non-zero component, in index order.
template <int Q, class R> __device__ __forceinline__ void rho_u(const R* f, R& rho, R& ux, R& uy, R& uz) {
    R r = f[0];
    static_for<Q - 1>([&](auto k) { r += f[decltype(k)::value + 1]; });
    r += R(1);
    R u[3] = {R(0), R(0), R(0)};
    static_for<3>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        R acc = R(0);
        bool first = true;
        static_for<(Q - 1) / 2>([&](auto p) {
            constexpr int i = 2 * decltype(p)::value + 1;
            constexpr int c = k == 0 ? VS<Q>::cx[i] : (k == 1 ? VS<Q>::cy[i] : VS<Q>::cz[i]);
            if (c != 0) {
                const R a = c > 0 ? f[i] : f[i + 1];
                const R b = c > 0 ? f[i + 1] : f[i];
                acc = first ? a - b : (acc + a) - b;
                first = false;
            }
        });
        u[k] = acc;
    });
    rho = r;
    ux = u[0] / r;
    uy = u[1] / r;
    uz = u[2] / r;
}

// calculate_forcing_terms (Guo), sim_kernels.cl:366-378. The unit direction
// components enter through exact sign selections.
template <int Q, class R> __device__ __forceinline__ void forcing(R ux, R uy, R uz, R fx, R fy, R fz, R* Fin) {
    R uF;
    if (VS<Q>::dims == 2) uF = -third<R>() * fmaR(ux, fx, uy * fy);
    else uF = -third<R>() * fmaR(ux, fx, fmaR(uy, fy, uz * fz));
    Fin[0] = (R(9) * weight_of_class<Q, R>(0)) * uF;
    static_for<Q - 1>([&](auto k) {
        constexpr int i = decltype(k)::value + 1;
        const R wi = weight<Q, R>(i);
        Fin[i] = (R(9) * wi) * fmaR(proj<Q, i, R>(fx, fy, fz), proj<Q, i, R>(ux, uy, uz) + third<R>(), uF);
    });
}

// calculate_a_eq (D3Q7), sim_kernels.cl:390-396
template <class R> __device__ __forceinline__ void a_eq(R Qv, R ux, R uy, R uz, R* qeq) {
    const R wsT4 = R(0.5) * Qv, wsTm1 = R(0.125) * (Qv - R(1));
    qeq[0] = fmaR(R(0.25), Qv, R(-0.25));
    qeq[1] = fmaR(wsT4, ux, wsTm1);
    qeq[2] = fmaR(wsT4, -ux, wsTm1);
    qeq[3] = fmaR(wsT4, uy, wsTm1);
    qeq[4] = fmaR(wsT4, -uy, wsTm1);
    qeq[5] = fmaR(wsT4, uz, wsTm1);
    qeq[6] = fmaR(wsT4, -uz, wsTm1);
}

// ---------------------------------------------------------------------------
// DDF storage formats (load/store macros, src/lbm/domain.rs:773-784)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t fp16c_encode(float x) {
    // custom 1-4-11 format, bias 15 (sim_kernels.cl:79-84). With b = bits +
    // 0x800 (the reference's rounding add) and e its exponent field:
    //  - e >= 113: code bits 14..0 = (b >> 12) & 0x7FFF (the reference's
    //    ((e - 112) << 11 & 0x7800) | m >> 12, since 112 is a multiple of 16);
    //  - e <= 112: the reference's ((M >> (124 - e)) + 1) >> 1 on the 24-bit
    //    significand M of x equals floor(|x| 2^25 + 1/2) (and 0 for e <= 100),
    //    computed as (floor(|x| 2^26) + 1) >> 1: the scaling is exact and the
    //    conversion truncates (adding 1/2 in float would round up at
    //    |x| 2^25 = 1/2 - 2^-25).
    // A NaN of either sign is encoded as the positive quiet NaN (0/0 on the GPU).
    const uint32_t b = (x != x ? 0x7FC00000u : __float_as_uint(x)) + 0x00000800u;
    const uint32_t nrm = (b >> 12) & 0x7FFFu;
    // floor(|x| 2^25 + 1/2): v_cvt_rpi_i32_f32 converts with round-half-up and
    // no intermediate addition (checked on the GPU for every float in [0, 4096):
    // equal to (u32(2y) + 1) >> 1), 2 operations instead of 4
    int den;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(den) : "v"(__builtin_fabsf(x) * 0x1p25f));
    // (b & 0x7FFFFFFF) >= 0x38800000 <=> |x| >= 0x387FF800 as a float, or x is NaN
    const bool big = !(__builtin_fabsf(x) < __uint_as_float(0x387FF800u));
    const uint32_t mag = big ? nrm : (uint32_t)den;
    return (uint16_t)(((b >> 16) & 0x8000u) | mag);
}
__device__ __forceinline__ float fp16c_decode(uint16_t x) {
    // sim_kernels.cl:85-90. Bits 14..0 of a code placed at bits 26..12 of a
    // float give the value times 2^-112 (normal codes as normal floats,
    // denormal codes as denormal floats, which the kernels keep), so one
    // exact multiplication by 2^112 decodes every code.
    const uint32_t u = (uint32_t)x;
    return __uint_as_float(((u & 0x7FFFu) << 12) | ((u & 0x8000u) << 16)) * 0x1p112f;
}

struct F32 {
    typedef float T;
    typedef float R;
    static __device__ __forceinline__ float ld(float v) { return v; }
    static __device__ __forceinline__ float st(float x) { return x; }
};
struct F16S {  // IEEE binary16 scaled by 2^15: vload_half*3.0517578E-5f / vstore_half_rte(x*32768.0f)
    typedef uint16_t T;
    typedef float R;
    static __device__ __forceinline__ float ld(uint16_t v) {
        return (float)__builtin_bit_cast(_Float16, v) * 3.0517578E-5f;
    }
    static __device__ __forceinline__ uint16_t st(float x) {
        const float y = x * 32768.0f;
        // the sign bit is re-applied because the compiler may fuse the scaling
        // into v_fma_mixlo_f16 (y*1 + 0), which turns -0 into +0
        const uint16_t h = __builtin_bit_cast(uint16_t, (_Float16)y) | (uint16_t)((__float_as_uint(y) >> 16) & 0x8000u);
        return y != y ? (uint16_t)0x7E00u : h;  // canonical NaN
    }
};
struct F16C {
    typedef uint16_t T;
    typedef float R;
    static __device__ __forceinline__ float ld(uint16_t v) { return fp16c_decode(v); }
    static __device__ __forceinline__ uint16_t st(float x) { return fp16c_encode(x); }
};

// FP64 instantiation: double DDFs and double arithmetic throughout
struct F64 {
    typedef double T;
    typedef double R;
    static __device__ __forceinline__ double ld(double v) { return v; }
    static __device__ __forceinline__ double st(double x) { return x; }
};

}  // namespace ion
