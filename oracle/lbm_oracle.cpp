/*
 * CPU oracle for the IonSolver extended-LBM MHD time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg. The product
 * library never links or calls it.
 *
 * A restatement of the reference kernel semantics, function by function;
 * every function cites the reference file:line it follows. Build flags:
 * -ffp-contract=off (no implicit fusing), so fma() appears exactly where the
 * reference source writes fma().
 */
#include "lbm_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

inline uint32_t bits_of(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
inline float float_of(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }

/* ------------------------------------------------------------------------ */
/* a1: storage codecs                                                       */
/* ------------------------------------------------------------------------ */

/* FP16C encode, sim_kernels.cl:79-84. Custom 1-4-11 format with exponent
 * bias 15 (range |x| < 2): add 0x800 below the kept mantissa, then truncate;
 * exponents 101..112 go to the denormal branch, smaller ones flush to +-0.
 * This is a bit-level format definition, restated as integer operations. */
uint16_t fp16c_encode(float x) {
    /* A NaN is encoded as the positive quiet NaN 0x7FC00000, which is what
     * 0/0 and inf-inf produce on GPUs; the sign of a NaN differs between
     * devices (x86 produces 0xFFC00000) and FP16C has no NaN code, so the
     * sign would otherwise select +1.5 or -1.5. */
    const uint32_t b = (x != x ? 0x7FC00000u : bits_of(x)) + 0x00000800u;
    const uint32_t e = (b & 0x7F800000u) >> 23;
    const uint32_t m = b & 0x007FFFFFu;
    uint32_t h = (b & 0x80000000u) >> 16;
    if (e > 112u) {
        h |= (((e - 112u) << 11) & 0x7800u) | (m >> 12);
    } else if (e > 100u) {
        h |= (((0x007FF800u + m) >> (124u - e)) + 1u) >> 1;
    }
    return (uint16_t)h;
}

/* FP16C decode, sim_kernels.cl:85-90. Normal codes: exponent e+112 in the
 * float; denormal codes (e == 0): value = m * 2^-25, exact in float. */
float fp16c_decode(uint16_t x) {
    const uint32_t e = (x & 0x7800u) >> 11;
    const uint32_t m = x & 0x07FFu;
    const float v = (e != 0u) ? float_of(((e + 112u) << 23) | (m << 12)) : (float)m * 0x1p-25f;
    return (x & 0x8000u) ? -v : v;
}

/* IEEE binary16 from binary32, round to nearest even (OpenCL
 * vstore_half_rte, used by the FP16S store macro, domain.rs:776). */
uint16_t half_encode_rte(float f) {
    const uint32_t x = bits_of(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7FFFFFFFu;
    if (ax > 0x7F800000u) return (uint16_t)0x7E00u; /* NaN: canonical positive quiet NaN */
    if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u); /* >= 65520: infinity */
    uint32_t q, rem, halfway;
    if (ax >= 0x38800000u) { /* normal binary16 */
        q = (((ax >> 23) - 112u) << 10) | ((ax >> 13) & 0x03FFu);
        rem = ax & 0x1FFFu;
        halfway = 0x1000u;
    } else {
        const uint32_t e = ax >> 23;
        if (e < 102u) return (uint16_t)sign; /* below half the smallest denormal */
        const uint32_t shift = 126u - e;     /* 14 .. 24 */
        const uint32_t m = (ax & 0x007FFFFFu) | 0x00800000u;
        q = m >> shift;
        rem = m & ((1u << shift) - 1u);
        halfway = 1u << (shift - 1u);
    }
    if (rem > halfway || (rem == halfway && (q & 1u))) q += 1u;
    return (uint16_t)(sign | q);
}

/* IEEE binary16 to binary32 (exact). */
float half_decode(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1Fu;
    const uint32_t m = h & 0x03FFu;
    if (e == 0u) {
        const float v = (float)m * 0x1p-24f;
        return sign ? -v : v;
    }
    if (e == 31u) return float_of(sign | 0x7F800000u | (m << 13));
    return float_of(sign | ((e + 112u) << 23) | (m << 13));
}

/* Storage policies: the load/store macros of domain.rs:773-784. */
struct StF32 {
    typedef float T; typedef float R;
    static inline R ld(T v) { return v; }
    static inline T st(R x) { return x; }
};
struct StF16S { /* vload_half(o,p)*3.0517578E-5f / vstore_half_rte((x)*32768.0f,o,p) */
    typedef uint16_t T; typedef float R;
    static inline R ld(T v) { return half_decode(v) * 3.0517578E-5f; }
    static inline T st(R x) { return half_encode_rte(x * 32768.0f); }
};
struct StF16C { /* half_to_float_custom / float_to_half_custom */
    typedef uint16_t T; typedef float R;
    static inline R ld(T v) { return fp16c_decode(v); }
    static inline T st(R x) { return fp16c_encode(x); }
};
struct StF64 { /* FP64 oracle: same algorithm in double */
    typedef double T; typedef double R;
    static inline R ld(T v) { return v; }
    static inline T st(R x) { return x; }
};

/* ------------------------------------------------------------------------ */
/* Velocity sets: direction tables in the reference ordering                */
/* (sim_kernels.cl:336-339 c(), 277-286 neighbors(), weights domain.rs:757-768) */
/* ------------------------------------------------------------------------ */

template <int Q> struct VSet;
template <> struct VSet<9> {
    static constexpr int dims = 2;
    static constexpr int c[3][9] = {
        {0, 1, -1, 0, 0, 1, -1, 1, -1},
        {0, 0, 0, 1, -1, 1, -1, -1, 1},
        {0, 0, 0, 0, 0, 0, 0, 0, 0}};
    /* weight class of each direction: 0 = W0, 1 = WS, 2 = WE/WC */
    static constexpr int wc[9] = {0, 1, 1, 1, 1, 2, 2, 2, 2};
    template <class R> static R weight(int k) {
        return k == 0 ? R(1) / R(2.25) : (k == 1 ? R(1) / R(9) : R(1) / R(36));
    }
};
template <> struct VSet<15> {
    static constexpr int dims = 3;
    static constexpr int c[3][15] = {
        {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1},
        {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1},
        {0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1, 1, -1}};
    static constexpr int wc[15] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2};
    template <class R> static R weight(int k) {
        return k == 0 ? R(1) / R(4.5) : (k == 1 ? R(1) / R(9) : R(1) / R(72));
    }
};
template <> struct VSet<19> {
    static constexpr int dims = 3;
    static constexpr int c[3][19] = {
        {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0},
        {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1},
        {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1}};
    static constexpr int wc[19] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
    template <class R> static R weight(int k) {
        return k == 0 ? R(1) / R(3) : (k == 1 ? R(1) / R(18) : R(1) / R(36));
    }
};
constexpr int VSet<9>::c[3][9];
constexpr int VSet<9>::wc[9];
constexpr int VSet<15>::c[3][15];
constexpr int VSet<15>::wc[15];
constexpr int VSet<19>::c[3][19];
constexpr int VSet<19>::wc[19];

/* "1/3" as written in the forcing terms (0.33333334f) and its FP64 value. */
template <class R> inline R third();
template <> inline float third<float>() { return 0.33333334f; }
template <> inline double third<double>() { return 1.0 / 3.0; }

/* Dot product of direction i with (X, Y, Z), accumulated left to right with
 * the reference's sign convention (first non-zero term carries its sign,
 * e.g. D3Q15 "u3 = -ux+uy+uz"), sim_kernels.cl:166-189. */
template <int Q, class R>
inline R proj(int i, R X, R Y, R Z) {
    const R v[3] = {X, Y, Z};
    R s = R(0);
    bool first = true;
    for (int k = 0; k < 3; k++) {
        const int ck = VSet<Q>::c[k][i];
        if (ck == 0) continue;
        if (first) { s = ck > 0 ? v[k] : -v[k]; first = false; }
        else { s = ck > 0 ? s + v[k] : s - v[k]; }
    }
    return s;
}

/* a4: calculate_f_eq, sim_kernels.cl:155-207. */
template <int Q, class R>
void f_eq(R rho, R ux, R uy, R uz, R* feq) {
    const R c3 = R(-3) * ((ux * ux + uy * uy) + uz * uz);
    const R rhom1 = rho - R(1);
    ux *= R(3);
    uy *= R(3);
    uz *= R(3);
    feq[0] = VSet<Q>::template weight<R>(0) * std::fma(rho, R(0.5) * c3, rhom1);
    const R ws = VSet<Q>::template weight<R>(1), we = VSet<Q>::template weight<R>(2);
    const R rhos = ws * rho, rhoe = we * rho, rhom1s = ws * rhom1, rhom1e = we * rhom1;
    for (int i = 1; i < Q; i += 2) {
        const bool straight = VSet<Q>::wc[i] == 1;
        const R rw = straight ? rhos : rhoe, rm = straight ? rhom1s : rhom1e;
        const R s = proj<Q, R>(i, ux, uy, uz);
        feq[i] = std::fma(rw, std::fma(R(0.5), std::fma(s, s, c3), s), rm);
        feq[i + 1] = std::fma(rw, std::fma(R(0.5), std::fma(s, s, c3), -s), rm);
    }
}

/* a4: calculate_rho_u, sim_kernels.cl:208-233. Density is summed in index
 * order and 1 is added last; each velocity component alternates + and -
 * over the direction pairs with a non-zero component, in index order. */
template <int Q, class R>
void rho_u(const R* f, R* rhon, R* uxn, R* uyn, R* uzn) {
    R rho = f[0];
    for (int i = 1; i < Q; i++) rho += f[i];
    rho += R(1);
    R u[3];
    for (int k = 0; k < 3; k++) {
        R acc = R(0);
        bool first = true;
        for (int i = 1; i < Q; i += 2) {
            const int ck = VSet<Q>::c[k][i];
            if (ck == 0) continue;
            const R a = ck > 0 ? f[i] : f[i + 1];
            const R b = ck > 0 ? f[i + 1] : f[i];
            if (first) { acc = a - b; first = false; }
            else { acc = (acc + a) - b; }
        }
        u[k] = acc;
    }
    *rhon = rho;
    *uxn = u[0] / rho;
    *uyn = u[1] / rho;
    *uzn = u[2] / rho;
}

/* a4: calculate_forcing_terms (Guo forcing), sim_kernels.cl:366-378. The
 * direction components are multiplied in literally, as the source does. */
template <int Q, class R>
void forcing(R ux, R uy, R uz, R fx, R fy, R fz, R* Fin) {
    R uF;
    if (VSet<Q>::dims == 2) uF = -third<R>() * std::fma(ux, fx, uy * fy);
    else uF = -third<R>() * std::fma(ux, fx, std::fma(uy, fy, uz * fz));
    Fin[0] = (R(9) * VSet<Q>::template weight<R>(0)) * uF;
    for (int i = 1; i < Q; i++) {
        const R cx = R(VSet<Q>::c[0][i]), cy = R(VSet<Q>::c[1][i]), cz = R(VSet<Q>::c[2][i]);
        const R wi = VSet<Q>::template weight<R>(VSet<Q>::wc[i]);
        Fin[i] = (R(9) * wi) * std::fma(cx * fx + cy * fy + cz * fz, cx * ux + cy * uy + cz * uz + third<R>(), uF);
    }
}

/* a4: calculate_a_eq (D3Q7, lattice speed of sound^2 = 1/4), sim_kernels.cl:390-396. */
template <class R>
void a_eq(R Qv, R ux, R uy, R uz, R* qeq) {
    const R wsT4 = R(0.5) * Qv, wsTm1 = R(0.125) * (Qv - R(1));
    qeq[0] = std::fma(R(0.25), Qv, R(-0.25));
    qeq[1] = std::fma(wsT4, ux, wsTm1);
    qeq[2] = std::fma(wsT4, -ux, wsTm1);
    qeq[3] = std::fma(wsT4, uy, wsTm1);
    qeq[4] = std::fma(wsT4, -uy, wsTm1);
    qeq[5] = std::fma(wsT4, uz, wsTm1);
    qeq[6] = std::fma(wsT4, -uz, wsTm1);
}

template <class R> inline R clampR(R x, R lo, R hi) { return std::fmin(std::fmax(x, lo), hi); }

/* Pairwise tree sum with +0 padding to the next power of two (the LOD
 * accumulation order, see lbm_oracle.h). */
template <class R>
R tree_sum(std::vector<R>& a, uint32_t n) {
    uint32_t P = 1;
    while (P < n) P <<= 1;
    a.resize(P);
    for (uint32_t i = n; i < P; i++) a[i] = R(0);
    for (uint32_t s = P >> 1; s >= 1; s >>= 1) {
        for (uint32_t i = 0; i < s; i++) a[i] = a[i] + a[i + s];
    }
    return a[0];
}

/* ------------------------------------------------------------------------ */
/* Simulation state                                                         */
/* ------------------------------------------------------------------------ */

struct SimBase {
    orc_params p{};
    uint64_t t = 0;
    bool initialized = false;
    virtual ~SimBase() {}
    virtual void* buffer(int id, uint64_t* count, uint32_t* esz) = 0;
    virtual void k_initialize() = 0;
    virtual void k_stream_collide() = 0;
    virtual void k_update_fields() = 0;
    virtual void k_clear_qu_lod() = 0;
    virtual void k_update_e_b() = 0;
    virtual double sum_rho() = 0;
    virtual double sum_q() = 0;
    virtual uint64_t n_lod() const = 0;
    bool mhd() const { return (p.ext & ORC_EXT_MAGNETO_HYDRO) != 0; }
    bool ecr() const { return mhd() && (p.ext & ORC_EXT_SUBGRID_ECR) != 0; }

    /* Lbm::initialize, src/lbm/mod.rs:214-231 (single domain: the
     * communicate_* calls are no-ops). */
    void initialize() {
        t = 1;
        k_initialize();
        if (mhd()) k_update_e_b();
        t = 0;
        initialized = true;
    }
    /* Lbm::do_time_step, src/lbm/mod.rs:250-272. */
    void do_time_step() {
        if (mhd()) k_clear_qu_lod();
        k_stream_collide();
        if (mhd()) k_update_e_b();
        t += 1;
    }
    /* Lbm::run, src/lbm/mod.rs:235-246. */
    void run(uint64_t steps) {
        if (!initialized) initialize();
        for (uint64_t i = 0; i < steps; i++) do_time_step();
    }
};

template <int Q, class S>
struct Sim : SimBase {
    typedef typename S::T T;
    typedef typename S::R R;
    uint32_t NX, NY, NZ;
    uint64_t N;
    uint32_t D = 0, nd = 1;
    uint64_t n_lod_own = 1;
    uint32_t bx = 1, by = 1, bz = 1; /* LOD block size */
    std::vector<T> fi, ei, fqi, eti;
    std::vector<R> rho, u, F, e_stat, b_stat, e_dyn, b_dyn, q, qu_lod, e_var, et;
    std::vector<uint8_t> flags;
    std::vector<R> lod_cell; /* per-cell LOD contributions (4 per cell) */
    R W, C, KE, KMU, KKGE, WQ, FX, FY, FZ, KKBME, KEABS;

    bool ext(uint32_t bit) const { return (p.ext & bit) != 0; }

    explicit Sim(const orc_params& pp) {
        p = pp;
        NX = p.nx; NY = p.ny; NZ = p.nz;
        N = (uint64_t)NX * NY * NZ;
        W = R(p.w); C = R(0.57735027f);
        KE = R(p.ke); KMU = R(p.kmu); KKGE = R(p.kkge); WQ = R(p.wq);
        FX = R(p.fx); FY = R(p.fy); FZ = R(p.fz);
        KKBME = R(p.kkbme); KEABS = R(p.keabs);
        fi.assign(N * Q, T(0));
        rho.assign(N, R(1));
        u.assign(3 * N, R(0));
        flags.assign(N, 0);
        if (ext(ORC_EXT_FORCE_FIELD)) F.assign(3 * N, R(0));
        if (mhd()) {
            D = p.lod_depth;
            nd = 1u << D;
            /* n_lod sizing, src/lbm/domain.rs:110-126 (3D, single domain) */
            n_lod_own = 1;
            for (uint32_t i = 0; i < D; i++) n_lod_own += (uint64_t)1 << (3 * (i + 1));
            e_stat.assign(3 * N, R(0)); b_stat.assign(3 * N, R(0));
            e_dyn.assign(3 * N, R(0)); b_dyn.assign(3 * N, R(0));
            fqi.assign(N * 7, T(0));
            ei.assign(N * Q, T(0));
            q.assign(N, R(0));
            qu_lod.assign(n_lod_own * 4, R(0));
            if (ecr()) {
                eti.assign(N * 7, T(0));
                et.assign(N, R(0));
                e_var.assign(3 * N, R(0));
            }
            if (D > 0) {
                bx = NX / nd; by = NY / nd; bz = NZ / nd;
                lod_cell.assign(N * 4, R(0));
            }
        }
    }

    uint64_t n_lod() const override { return mhd() ? n_lod_own : 0; }

    void* buffer(int id, uint64_t* count, uint32_t* esz) override {
        auto put = [&](void* ptr, uint64_t c, uint32_t e) -> void* {
            if (count) *count = c;
            if (esz) *esz = e;
            return c ? ptr : nullptr;
        };
        switch (id) {
            case ORC_BUF_FI: return put(fi.data(), fi.size(), sizeof(T));
            case ORC_BUF_RHO: return put(rho.data(), rho.size(), sizeof(R));
            case ORC_BUF_U: return put(u.data(), u.size(), sizeof(R));
            case ORC_BUF_FLAGS: return put(flags.data(), flags.size(), 1);
            case ORC_BUF_F: return put(F.data(), F.size(), sizeof(R));
            case ORC_BUF_E_STAT: return put(e_stat.data(), e_stat.size(), sizeof(R));
            case ORC_BUF_B_STAT: return put(b_stat.data(), b_stat.size(), sizeof(R));
            case ORC_BUF_E_DYN: return put(e_dyn.data(), e_dyn.size(), sizeof(R));
            case ORC_BUF_B_DYN: return put(b_dyn.data(), b_dyn.size(), sizeof(R));
            case ORC_BUF_FQI: return put(fqi.data(), fqi.size(), sizeof(T));
            case ORC_BUF_Q: return put(q.data(), q.size(), sizeof(R));
            case ORC_BUF_QU_LOD: return put(qu_lod.data(), qu_lod.size(), sizeof(R));
            case ORC_BUF_EI: return put(ei.data(), ei.size(), sizeof(T));
            case ORC_BUF_E_VAR: return put(e_var.data(), e_var.size(), sizeof(R));
            case ORC_BUF_ETI: return put(eti.data(), eti.size(), sizeof(T));
            case ORC_BUF_ET: return put(et.data(), et.size(), sizeof(R));
        }
        return put(nullptr, 0, 0);
    }

    /* a2: coordinates / index_f / neighbors, sim_kernels.cl:134-154, 248-302. */
    inline void coords(uint32_t n, uint32_t& x, uint32_t& y, uint32_t& z) const {
        const uint32_t t2 = n % (NX * NY);
        x = t2 % NX; y = t2 / NX; z = n / (NX * NY);
    }
    inline uint64_t idx(uint32_t n, uint32_t i) const { return (uint64_t)i * N + (uint64_t)n; }
    void neighbors(uint32_t n, uint32_t* j) const {
        uint32_t x, y, z;
        coords(n, x, y, z);
        const uint32_t xs[3] = {x, (x + 1u) % NX, (x + NX - 1u) % NX};
        const uint32_t ys[3] = {y * NX, ((y + 1u) % NY) * NX, ((y + NY - 1u) % NY) * NX};
        const uint32_t zs[3] = {z * NY * NX, ((z + 1u) % NZ) * NY * NX, ((z + NZ - 1u) % NZ) * NY * NX};
        for (int i = 0; i < Q; i++) {
            const int cx = VSet<Q>::c[0][i], cy = VSet<Q>::c[1][i], cz = VSet<Q>::c[2][i];
            j[i] = xs[cx == 0 ? 0 : (cx > 0 ? 1 : 2)] + ys[cy == 0 ? 0 : (cy > 0 ? 1 : 2)] +
                   zs[cz == 0 ? 0 : (cz > 0 ? 1 : 2)];
        }
    }
    /* neighbors_a, sim_kernels.cl:382-389: the first 7 D3Q19 directions. */
    void neighbors7(uint32_t n, uint32_t* j7) const {
        uint32_t x, y, z;
        coords(n, x, y, z);
        const uint32_t x0 = x, xp = (x + 1u) % NX, xm = (x + NX - 1u) % NX;
        const uint32_t y0 = y * NX, yp = ((y + 1u) % NY) * NX, ym = ((y + NY - 1u) % NY) * NX;
        const uint32_t z0 = z * NY * NX, zp = ((z + 1u) % NZ) * NY * NX, zm = ((z + NZ - 1u) % NZ) * NY * NX;
        j7[0] = n;
        j7[1] = xp + y0 + z0; j7[2] = xm + y0 + z0;
        j7[3] = x0 + yp + z0; j7[4] = x0 + ym + z0;
        j7[5] = x0 + y0 + zp; j7[6] = x0 + y0 + zm;
    }

    /* a3: esoteric pull, load_f/store_f sim_kernels.cl:234-247 and
     * load_a/store_a :397-410 (NQ = number of directions of the lattice). */
    template <int NQ>
    void load(uint32_t n, R* fh, const T* buf, const uint32_t* j, uint64_t tt) const {
        fh[0] = S::ld(buf[idx(n, 0)]);
        for (uint32_t i = 1; i < (uint32_t)NQ; i += 2) {
            fh[i] = S::ld(buf[idx(n, (tt % 2) ? i : i + 1)]);
            fh[i + 1] = S::ld(buf[idx(j[i], (tt % 2) ? i + 1 : i)]);
        }
    }
    template <int NQ>
    void store(uint32_t n, const R* fh, T* buf, const uint32_t* j, uint64_t tt) const {
        buf[idx(n, 0)] = S::st(fh[0]);
        for (uint32_t i = 1; i < (uint32_t)NQ; i += 2) {
            buf[idx(j[i], (tt % 2) ? i + 1 : i)] = S::st(fh[i]);
            buf[idx(n, (tt % 2) ? i : i + 1)] = S::st(fh[i + 1]);
        }
    }

    /* lod_index, sim_kernels.cl:425-433 */
    uint32_t lod_index(uint32_t n, uint32_t d) const {
        uint32_t x, y, z;
        coords(n, x, y, z);
        const uint32_t ndd = 1u << d;
        return x / (NX / ndd) + (y / (NY / ndd) + z / (NZ / ndd) * ndd) * ndd;
    }

    /* a6: initialize, sim_kernels.cl:760-832 (t = 1 layout). */
    void k_initialize() override {
        const uint64_t tt = 1;
#pragma omp parallel for schedule(static)
        for (int64_t nn = 0; nn < (int64_t)N; nn++) {
            const uint32_t n = (uint32_t)nn;
            const uint8_t bo = flags[n] & 0x1F;
            uint32_t j[Q];
            neighbors(n, j);
            if (bo == 0x01) { /* solid: zero velocity and charge */
                u[n] = R(0); u[N + n] = R(0); u[2 * N + n] = R(0);
                if (mhd()) q[n] = R(0);
            }
            R feq[Q];
            f_eq<Q, R>(rho[n], u[n], u[N + n], u[2 * N + n], feq);
            store<Q>(n, feq, fi.data(), j, tt);
            if (mhd()) {
                R qeq[7];
                a_eq<R>(q[n], u[n], u[N + n], u[2 * N + n], qeq);
                uint32_t j7[7];
                neighbors7(n, j7);
                store<7>(n, qeq, fqi.data(), j7, tt);
                for (int k = 0; k < 3; k++) {
                    b_dyn[k * N + n] = b_stat[k * N + n];
                    e_dyn[k * N + n] = e_stat[k * N + n];
                }
                /* density 0 (:823-825), or the build extension's ei_rho0 off solid cells */
                f_eq<Q, R>(bo == 0x01 ? R(0) : R(p.ei_rho0), u[n], u[N + n], u[2 * N + n], feq);
                store<Q>(n, feq, ei.data(), j, tt);
                if (ecr()) { /* electron temperature DDFs (:827-831) */
                    a_eq<R>(et[n], u[n], u[N + n], u[2 * N + n], qeq);
                    store<7>(n, qeq, eti.data(), j7, tt);
                }
            }
        }
    }

    /* length(v) as sqrt(dot(v, v)), dot as (x x + y y) + z z */
    static R len3(R x, R y, R z) { return std::sqrt((x * x + y * y) + z * z); }

    /* grad_mag_v(n, B_dyn), sim_kernels.cl:453-461: zero on the lattice
     * faces, else per axis mag(c + e) - mag(c - e) / 2 (as written: only the
     * second term is halved) */
    void grad_mag_b(uint32_t n, R* g) const {
        const uint32_t x = n % NX, y = (n / NX) % NY, z = (uint32_t)(n / ((uint64_t)NX * NY));
        if (x == 0 || x == NX - 1 || y == 0 || y == NY - 1 || z == 0 || z == NZ - 1) {
            g[0] = g[1] = g[2] = R(0);
            return;
        }
        auto mag = [&](uint64_t i) { return len3(b_dyn[i], b_dyn[N + i], b_dyn[2 * N + i]); };
        const uint64_t sy = NX, sz = (uint64_t)NX * NY;
        g[0] = mag(n + 1) - mag(n - 1) / R(2);
        g[1] = mag(n + sy) - mag(n - sy) / R(2);
        g[2] = mag(n + sz) - mag(n - sz) / R(2);
    }

    /* SUBGRID_ECR for one cell, sim_kernels.cl:557-629: electron temperature
     * streaming, ECR heating by the components of E_var normal to B, the
     * guiding-centre drift along -grad|B|, temperature collision, ionisation.
     * The cyclotron frequency and relative absorption of :582-585 are
     * computed there but never used, so they are not restated. */
    void ecr_cell(uint32_t n, const uint32_t* j7, uint64_t tt, R Bx, R By, R Bz, R& rhon_e, R& uxn_e, R& uyn_e,
                  R& uzn_e, R& rhon_q) {
        R ethn[7];
        load<7>(n, ethn, eti.data(), j7, tt);
        R Etn = R(0);
        for (int i = 0; i < 7; i++) Etn += ethn[i];
        Etn += R(1);
        const R Evx = e_var[n], Evy = e_var[N + n], Evz = e_var[2 * N + n];
        const R lb = len3(Bx, By, Bz);
        const R s = ((Evx * Bx + Evy * By) + Evz * Bz) / (lb * lb);
        const R env = len3(Evx - s * Bx, Evy - s * By, Evz - s * Bz);
        Etn += KEABS / KKBME * (rhon_e + R(0.00001f)) / KKGE * (env * env);
        R g[3];
        grad_mag_b(n, g);
        const R kt = KKBME * Etn;
        const R dx = kt * g[0] / lb, dy = kt * g[1] / lb, dz = kt * g[2] / lb;
        uxn_e += dx;
        uyn_e += dy;
        uzn_e += dz;
        Etn -= len3(dx, dy, dz) / KKBME;
        R eteq[7];
        a_eq<R>(Etn, uxn_e, uyn_e, uzn_e, eteq);
        if (ext(ORC_EXT_UPDATE_FIELDS)) et[n] = Etn;
        for (int i = 0; i < 7; i++) ethn[i] = std::fma(R(1) - WQ, ethn[i], WQ * eteq[i]);
        store<7>(n, ethn, eti.data(), j7, tt);
        const R dq = R(0.0001f) * Etn;
        rhon_e += dq;
        rhon_q += dq;
    }

    /* a5: stream_collide for one cell, sim_kernels.cl:465-758. */
    void stream_collide_cell(uint32_t n, uint64_t tt) {
        const uint8_t flagsn = flags[n];
        const uint8_t bo = flagsn & 0x1F;
        if (bo == 0x01) return;
        const bool eqb = ext(ORC_EXT_EQUILIBRIUM_BOUNDARIES);
        const bool vforce = ext(ORC_EXT_VOLUME_FORCE);
        const bool is_e = eqb && bo == 0x02;

        uint32_t j[Q];
        neighbors(n, j);
        R fhn[Q];
        load<Q>(n, fhn, fi.data(), j, tt);
        const uint64_t nxi = n, nyi = N + n, nzi = 2 * N + n;

        R rhon, uxn, uyn, uzn;
        if (is_e) {
            rhon = rho[n]; uxn = u[nxi]; uyn = u[nyi]; uzn = u[nzi];
        } else {
            rho_u<Q, R>(fhn, &rhon, &uxn, &uyn, &uzn);
        }
        R fxn = FX, fyn = FY, fzn = FZ;
        R Fin[Q], feq[Q];
        const R w = W;
        const R c_tau = std::fma(w, R(-0.5), R(1));
        if (ext(ORC_EXT_FORCE_FIELD)) {
            fxn += F[nxi]; fyn += F[nyi]; fzn += F[nzi];
        }
        if (mhd()) {
            const R Bx = b_dyn[nxi], By = b_dyn[nyi], Bz = b_dyn[nzi];
            const R Ex = e_dyn[nxi], Ey = e_dyn[nyi], Ez = e_dyn[nzi];
            /* electron gas part 1 (:537-540) */
            R ehn[Q];
            load<Q>(n, ehn, ei.data(), j, tt);
            R rhon_e, uxn_e, uyn_e, uzn_e;
            rho_u<Q, R>(ehn, &rhon_e, &uxn_e, &uyn_e, &uzn_e);
            /* charge advection part 1 (:547-553) */
            uint32_t j7[7];
            neighbors7(n, j7);
            R qhn[7];
            load<7>(n, qhn, fqi.data(), j7, tt);
            R rhon_q = R(0);
            for (int i = 0; i < 7; i++) rhon_q += qhn[i];
            rhon_q += R(1);
            if (ecr()) ecr_cell(n, j7, tt, Bx, By, Bz, rhon_e, uxn_e, uyn_e, uzn_e, rhon_q);
            /* charge inlet (ORC_EXT_CHARGE_INLET): the cell's charge is held */
            const bool inlet = is_e && ext(ORC_EXT_CHARGE_INLET);
            if (inlet) rhon_q = q[n];
            /* charge advection part 2 (:633-637) */
            else q[n] = rhon_q - rhon_e;
            R qeq[7];
            a_eq<R>(rhon_q, uxn, uyn, uzn, qeq);
            for (int i = 0; i < 7; i++) qhn[i] = inlet ? qeq[i] : std::fma(R(1) - WQ, qhn[i], WQ * qeq[i]);
            store<7>(n, qhn, fqi.data(), j7, tt);
            /* electron gas part 2 (:641-656): F = -rho_e (E + u_e x B) */
            const R crx = uyn_e * Bz - uzn_e * By;
            const R cry = uzn_e * Bx - uxn_e * Bz;
            const R crz = uxn_e * By - uyn_e * Bx;
            const R efx = -rhon_e * (Ex + crx), efy = -rhon_e * (Ey + cry), efz = -rhon_e * (Ez + crz);
            const R rho2_e = R(0.5) / (rhon_e * KKGE);
            uxn_e = clampR(std::fma(efx, rho2_e, uxn_e), -C, C);
            uyn_e = clampR(std::fma(efy, rho2_e, uyn_e), -C, C);
            uzn_e = clampR(std::fma(efz, rho2_e, uzn_e), -C, C);
            forcing<Q, R>(uxn_e, uyn_e, uzn_e, efx, efy, efz, Fin);
            f_eq<Q, R>(rhon_e, uxn_e, uyn_e, uzn_e, feq);
            for (int i = 0; i < Q; i++) Fin[i] *= c_tau;
            for (int i = 0; i < Q; i++)
                ehn[i] = is_e ? feq[i] : std::fma(R(1) - w, ehn[i], std::fma(w, feq[i], Fin[i]));
            store<Q>(n, ehn, ei.data(), j, tt);
            /* EM force on the gas (:660-662) */
            fxn += rhon_q * (Ex + uyn * Bz - uzn * By);
            fyn += rhon_q * (Ey + uzn * Bx - uxn * Bz);
            fzn += rhon_q * (Ez + uxn * By - uyn * Bx);
            /* LOD contribution (:666-677), summed later in tree order */
            if (D > 0) {
                const R ils = R(1) / R((float)((uint64_t)bx * by * bz));
                lod_cell[4 * (uint64_t)n + 0] = rhon_q - rhon_e;
                lod_cell[4 * (uint64_t)n + 1] = uxn * ils;
                lod_cell[4 * (uint64_t)n + 2] = uyn * ils;
                lod_cell[4 * (uint64_t)n + 3] = uzn * ils;
            }
        }
        if (vforce) {
            const R rho2 = R(0.5) / rhon;
            uxn = clampR(std::fma(fxn, rho2, uxn), -C, C);
            uyn = clampR(std::fma(fyn, rho2, uyn), -C, C);
            uzn = clampR(std::fma(fzn, rho2, uzn), -C, C);
            forcing<Q, R>(uxn, uyn, uzn, fxn, fyn, fzn, Fin);
        } else {
            uxn = clampR(uxn, -C, C);
            uyn = clampR(uyn, -C, C);
            uzn = clampR(uzn, -C, C);
            for (int i = 0; i < Q; i++) Fin[i] = R(0);
        }
        if (ext(ORC_EXT_UPDATE_FIELDS) && !is_e) {
            rho[n] = rhon; u[nxi] = uxn; u[nyi] = uyn; u[nzi] = uzn;
        }
        f_eq<Q, R>(rhon, uxn, uyn, uzn, feq);
        if (p.relaxation_time == 0) { /* SRT (:714-723) */
            if (vforce)
                for (int i = 0; i < Q; i++) Fin[i] *= c_tau;
            for (int i = 0; i < Q; i++)
                fhn[i] = is_e ? feq[i] : std::fma(R(1) - w, fhn[i], std::fma(w, feq[i], Fin[i]));
        } else { /* TRT (:725-755) */
            const R wp = w;
            const R wm = R(1) / (R(0.1875) / (R(1) / w - R(0.5)) + R(0.5));
            if (vforce) {
                const R c_taup = std::fma(wp, R(-0.25), R(0.5)), c_taum = std::fma(wm, R(-0.25), R(0.5));
                R Fib[Q];
                Fib[0] = Fin[0];
                for (int i = 1; i < Q; i += 2) { Fib[i] = Fin[i + 1]; Fib[i + 1] = Fin[i]; }
                for (int i = 0; i < Q; i++)
                    Fin[i] = std::fma(c_taup, Fin[i] + Fib[i], c_taum * (Fin[i] - Fib[i]));
            }
            R fhb[Q], feb[Q];
            fhb[0] = fhn[0];
            feb[0] = feq[0];
            for (int i = 1; i < Q; i += 2) {
                fhb[i] = fhn[i + 1]; fhb[i + 1] = fhn[i];
                feb[i] = feq[i + 1]; feb[i + 1] = feq[i];
            }
            for (int i = 0; i < Q; i++)
                fhn[i] = is_e ? feq[i]
                              : std::fma(R(0.5) * wp, feq[i] - fhn[i] + feb[i] - fhb[i],
                                         std::fma(R(0.5) * wm, feq[i] - feb[i] - fhn[i] + fhb[i], fhn[i] + Fin[i]));
        }
        store<Q>(n, fhn, fi.data(), j, tt);
    }

    void k_stream_collide() override {
        const uint64_t tt = t;
        const bool lod = mhd() && D > 0;
        if (lod) std::fill(lod_cell.begin(), lod_cell.end(), R(0));
#pragma omp parallel for schedule(static)
        for (int64_t nn = 0; nn < (int64_t)N; nn++) stream_collide_cell((uint32_t)nn, tt);
        if (lod) lod_reduce();
    }

    /* Level-of-detail sums in the fixed tree order (lbm_oracle.h). Writes
     * the top-level LODs at offset 0 (single domain, sim_kernels.cl:667-671). */
    void lod_reduce() {
        const uint32_t nseg_x = nd; /* row segments per row */
        std::vector<R> seg((uint64_t)NZ * NY * nseg_x * 4);
#pragma omp parallel
        {
            std::vector<R> a;
#pragma omp for schedule(static)
            for (int64_t row = 0; row < (int64_t)NZ * NY; row++) {
                for (uint32_t X = 0; X < nseg_x; X++) {
                    for (int k = 0; k < 4; k++) {
                        a.assign(bx, R(0));
                        for (uint32_t i = 0; i < bx; i++)
                            a[i] = lod_cell[4 * ((uint64_t)row * NX + X * bx + i) + k];
                        seg[4 * ((uint64_t)row * nseg_x + X) + k] = tree_sum(a, bx);
                    }
                }
            }
        }
        const uint64_t nlod = (uint64_t)nd * nd * nd;
#pragma omp parallel
        {
            std::vector<R> a;
#pragma omp for schedule(static)
            for (int64_t L = 0; L < (int64_t)nlod; L++) {
                const uint32_t X = (uint32_t)(L % nd), Y = (uint32_t)((L / nd) % nd), Z = (uint32_t)(L / ((uint64_t)nd * nd));
                for (int k = 0; k < 4; k++) {
                    a.assign((uint64_t)by * bz, R(0));
                    for (uint32_t zz = 0; zz < bz; zz++)
                        for (uint32_t yy = 0; yy < by; yy++) {
                            const uint64_t row = (uint64_t)(Z * bz + zz) * NY + (Y * by + yy);
                            a[yy + (uint64_t)zz * by] = seg[4 * (row * nseg_x + X) + k];
                        }
                    qu_lod[4 * L + k] = tree_sum(a, by * bz);
                }
            }
        }
    }

    /* update_fields, sim_kernels.cl:834-859 (not on the hot path). */
    void k_update_fields() override {
        const uint64_t tt = t;
#pragma omp parallel for schedule(static)
        for (int64_t nn = 0; nn < (int64_t)N; nn++) {
            const uint32_t n = (uint32_t)nn;
            if ((flags[n] & 0x1F) == 0x01) continue;
            uint32_t j[Q];
            neighbors(n, j);
            R fhn[Q];
            load<Q>(n, fhn, fi.data(), j, tt);
            R rhon, uxn, uyn, uzn;
            rho_u<Q, R>(fhn, &rhon, &uxn, &uyn, &uzn);
            rho[n] = rhon;
            u[n] = clampR(uxn, -C, C);
            u[N + n] = clampR(uyn, -C, C);
            u[2 * N + n] = clampR(uzn, -C, C);
        }
    }

    /* clear_qu_lod, sim_kernels.cl:995-1003 */
    void k_clear_qu_lod() override { std::fill(qu_lod.begin(), qu_lod.end(), R(0)); }

    /* a8: update_e_b_dynamic, sim_kernels.cl:897-993 (single domain). */
    void k_update_e_b() override {
        const uint32_t ndd = 1u << D;
        /* near-field block side: imax(N / (1<<nd), 1); OpenCL masks the shift
         * count to 5 bits for a 32-bit int. */
        const uint32_t sh = 1u << (ndd & 31u);
        const uint32_t dsx = std::max<int32_t>((int32_t)(NX / sh), 1);
        const uint32_t dsy = std::max<int32_t>((int32_t)(NY / sh), 1);
        const uint32_t dsz = std::max<int32_t>((int32_t)(NZ / sh), 1);
        const uint64_t n_top = (uint64_t)ndd * ndd * ndd;
        const uint64_t lo = n_lod_own - n_top, hi = n_lod_own;
        const R dcx = R((float)(NX / ndd)), dcy = R((float)(NY / ndd)), dcz = R((float)(NZ / ndd));
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t nn = 0; nn < (int64_t)N; nn++) {
            const uint32_t n = (uint32_t)nn;
            if ((flags[n] & 0x1F) == 0x01) continue;
            uint32_t cx, cy, cz;
            coords(n, cx, cy, cz);
            const R px = R((float)cx), py = R((float)cy), pz = R((float)cz);
            R ex = 0, ey = 0, ez = 0, bxx = 0, byy = 0, bzz = 0;
            auto add = [&](R qc, R vx, R vy, R vz, R rx, R ry, R rz) {
                const R s = std::sqrt((rx * rx + ry * ry) + rz * rz);
                const R cbm = (s * s) * s;
                const R prx = rx / cbm, pry = ry / cbm, prz = rz / cbm;
                ex = ex + qc * prx; ey = ey + qc * pry; ez = ez + qc * prz;
                const R crx = vy * prz - vz * pry, cry = vz * prx - vx * prz, crz = vx * pry - vy * prx;
                bxx = bxx + qc * crx; byy = byy + qc * cry; bzz = bzz + qc * crz;
            };
            /* near field: individual cells of the own block (:907-938) */
            const uint32_t xu = std::min((cx / dsx) * dsx + dsx, NX);
            const uint32_t yu = std::min((cy / dsy) * dsy + dsy, NY);
            const uint32_t zu = std::min((cz / dsz) * dsz + dsz, NZ);
            for (uint32_t x = (cx / dsx) * dsx; x < xu; x++)
                for (uint32_t y = (cy / dsy) * dsy; y < yu; y++)
                    for (uint32_t z = (cz / dsz) * dsz; z < zu; z++) {
                        const uint32_t nc = x + (y + z * NY) * NX;
                        if (nc == n) continue;
                        const R qc = q[nc];
                        if (qc == R(0)) continue;
                        add(qc, u[nc], u[(uint64_t)nc + N], u[(uint64_t)nc + 2 * N],
                            px - R((float)x), py - R((float)y), pz - R((float)z));
                    }
            /* medium distance: the own-domain LOD range (:941-955) */
            const uint32_t ndi = lod_index(n, D);
            for (uint64_t d = lo; d < hi; d++) {
                if (d == ndi) continue;
                const uint64_t t2 = d % ((uint64_t)ndd * ndd);
                const R lx = R((float)(t2 % ndd)) * dcx + R(0.5) * dcx;
                const R ly = R((float)(t2 / ndd)) * dcy + R(0.5) * dcy;
                const R lz = R((float)(d / ((uint64_t)ndd * ndd))) * dcz + R(0.5) * dcz;
                add(qu_lod[d * 4 + 0], qu_lod[d * 4 + 1], qu_lod[d * 4 + 2], qu_lod[d * 4 + 3],
                    px - lx, py - ly, pz - lz);
            }
            e_dyn[n] = e_stat[n] + KE * ex;
            e_dyn[N + n] = e_stat[N + n] + KE * ey;
            e_dyn[2 * N + n] = e_stat[2 * N + n] + KE * ez;
            b_dyn[n] = b_stat[n] + KMU * bxx;
            b_dyn[N + n] = b_stat[N + n] + KMU * byy;
            b_dyn[2 * N + n] = b_stat[2 * N + n] + KMU * bzz;
        }
    }

    double sum_rho() override {
        double s = 0;
#pragma omp parallel for reduction(+ : s)
        for (int64_t n = 0; n < (int64_t)N; n++)
            if ((flags[n] & 0x1F) != 0x01) s += (double)rho[n];
        return s;
    }
    double sum_q() override {
        double s = 0;
        if (!mhd()) return 0;
#pragma omp parallel for reduction(+ : s)
        for (int64_t n = 0; n < (int64_t)N; n++)
            if ((flags[n] & 0x1F) != 0x01) s += (double)q[n];
        return s;
    }
};

template <int Q>
SimBase* make_for_store(const orc_params& p) {
    switch (p.float_type) {
        case ORC_FP32: return new Sim<Q, StF32>(p);
        case ORC_FP16S: return new Sim<Q, StF16S>(p);
        case ORC_FP16C: return new Sim<Q, StF16C>(p);
        case ORC_FP64: return new Sim<Q, StF64>(p);
    }
    return nullptr;
}

}  // namespace

struct orc_sim {
    SimBase* s;
};

extern "C" {

orc_sim* orc_create(const orc_params* p, char* err) {
    auto fail = [&](const char* m) -> orc_sim* {
        if (err) std::snprintf(err, 256, "%s", m);
        return nullptr;
    };
    if (!p) return fail("null params");
    if (p->nx == 0 || p->ny == 0 || p->nz == 0) return fail("zero lattice size");
    const uint64_t N = (uint64_t)p->nx * p->ny * p->nz;
    if (N > 0xFFFFFFFFull) return fail("more than 2^32 cells");
    if (p->float_type < ORC_FP16S || p->float_type > ORC_FP64) return fail("bad float_type");
    if (p->velocity_set == ORC_D3Q27) return fail("D3Q27 is not supported (reference defines no DEF_WC for it)");
    if (p->velocity_set < ORC_D2Q9 || p->velocity_set > ORC_D3Q27) return fail("bad velocity_set");
    if (p->ext & ORC_EXT_MAGNETO_HYDRO) {
        if (p->velocity_set == ORC_D2Q9) return fail("magneto_hydro needs a 3D velocity set");
        if (!(p->ext & ORC_EXT_VOLUME_FORCE)) return fail("magneto_hydro needs volume_force");
        if (p->lod_depth > 7) return fail("lod_depth > 7");
        if (p->lod_depth > 0) {
            const uint32_t nd = 1u << p->lod_depth;
            if (p->nx % nd || p->ny % nd || p->nz % nd) return fail("lattice size not divisible by 2^lod_depth");
        }
    }
    if (p->ext & ORC_EXT_FORCE_FIELD && !(p->ext & ORC_EXT_VOLUME_FORCE)) return fail("force_field needs volume_force");
    SimBase* s = nullptr;
    switch (p->velocity_set) {
        case ORC_D2Q9: s = make_for_store<9>(*p); break;
        case ORC_D3Q15: s = make_for_store<15>(*p); break;
        case ORC_D3Q19: s = make_for_store<19>(*p); break;
    }
    if (!s) return fail("unsupported combination");
    return new orc_sim{s};
}

void orc_destroy(orc_sim* s) {
    if (s) { delete s->s; delete s; }
}
void* orc_buffer(orc_sim* s, int id, uint64_t* count, uint32_t* esz) { return s->s->buffer(id, count, esz); }
uint64_t orc_time_step(const orc_sim* s) { return s->s->t; }
void orc_set_time_step(orc_sim* s, uint64_t t) { s->s->t = t; }
uint64_t orc_n_lod(const orc_sim* s) { return s->s->n_lod(); }
void orc_initialize(orc_sim* s) { s->s->initialize(); }
void orc_run(orc_sim* s, uint64_t steps) { s->s->run(steps); }
void orc_do_time_step(orc_sim* s) { s->s->do_time_step(); }
void orc_kernel_initialize(orc_sim* s) { s->s->k_initialize(); }
void orc_kernel_stream_collide(orc_sim* s) { s->s->k_stream_collide(); }
void orc_kernel_update_fields(orc_sim* s) { s->s->k_update_fields(); }
void orc_kernel_clear_qu_lod(orc_sim* s) { s->s->k_clear_qu_lod(); }
void orc_kernel_update_e_b_dynamic(orc_sim* s) { s->s->k_update_e_b(); }
double orc_sum_rho(orc_sim* s) { return s->s->sum_rho(); }
double orc_sum_q(orc_sim* s) { return s->s->sum_q(); }

void orc_fp16c_encode(const float* x, uint16_t* h, uint64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) h[i] = fp16c_encode(x[i]);
}
void orc_fp16c_decode(const uint16_t* h, float* x, uint64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) x[i] = fp16c_decode(h[i]);
}
void orc_half_encode_rte(const float* x, uint16_t* h, uint64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) h[i] = half_encode_rte(x[i]);
}
void orc_half_decode(const uint16_t* h, float* x, uint64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) x[i] = half_decode(h[i]);
}
void orc_fp16s_store(const float* x, uint16_t* h, uint64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) h[i] = StF16S::st(x[i]);
}
void orc_fp16s_load(const uint16_t* h, float* x, uint64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) x[i] = StF16S::ld(h[i]);
}

void orc_f_eq(int vs, float rho, float ux, float uy, float uz, float* feq) {
    if (vs == ORC_D2Q9) f_eq<9, float>(rho, ux, uy, uz, feq);
    else if (vs == ORC_D3Q15) f_eq<15, float>(rho, ux, uy, uz, feq);
    else f_eq<19, float>(rho, ux, uy, uz, feq);
}
void orc_f_eq_f64(int vs, double rho, double ux, double uy, double uz, double* feq) {
    if (vs == ORC_D2Q9) f_eq<9, double>(rho, ux, uy, uz, feq);
    else if (vs == ORC_D3Q15) f_eq<15, double>(rho, ux, uy, uz, feq);
    else f_eq<19, double>(rho, ux, uy, uz, feq);
}
void orc_rho_u(int vs, const float* f, float* r) {
    if (vs == ORC_D2Q9) rho_u<9, float>(f, r, r + 1, r + 2, r + 3);
    else if (vs == ORC_D3Q15) rho_u<15, float>(f, r, r + 1, r + 2, r + 3);
    else rho_u<19, float>(f, r, r + 1, r + 2, r + 3);
}
void orc_rho_u_f64(int vs, const double* f, double* r) {
    if (vs == ORC_D2Q9) rho_u<9, double>(f, r, r + 1, r + 2, r + 3);
    else if (vs == ORC_D3Q15) rho_u<15, double>(f, r, r + 1, r + 2, r + 3);
    else rho_u<19, double>(f, r, r + 1, r + 2, r + 3);
}
void orc_forcing(int vs, float ux, float uy, float uz, float fx, float fy, float fz, float* fin) {
    if (vs == ORC_D2Q9) forcing<9, float>(ux, uy, uz, fx, fy, fz, fin);
    else if (vs == ORC_D3Q15) forcing<15, float>(ux, uy, uz, fx, fy, fz, fin);
    else forcing<19, float>(ux, uy, uz, fx, fy, fz, fin);
}
void orc_a_eq(float qv, float ux, float uy, float uz, float* qeq) { a_eq<float>(qv, ux, uy, uz, qeq); }

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_num_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
