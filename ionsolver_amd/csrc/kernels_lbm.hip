// gfx950 kernels of the lattice part of the time step: initialize,
// stream_collide (fluid + electron gas + D3Q7 charge advection + Lorentz
// force + level-of-detail charge sums) and update_fields.
//
// One work-item per cell, DDFs in the reference's SoA layout
// (index_f(n, i) = i*N + n, sim_kernels.cl:152-154), streamed in place by
// the esoteric-pull scheme (sim_kernels.cl:234-247), so every DDF slot is
// read once and written once per step by exactly one work-item.
#include <hip/hip_runtime.h>

#include "device_lattice.hpp"
#include "kernels.hpp"

namespace ion {

struct Nbr {
    uint32_t x0, xp, xm, y0, yp, ym, z0, zp, zm;
};

// calculate_indices, sim_kernels.cl:248-259 (periodic wrap)
__device__ __forceinline__ Nbr neighbours(uint32_t n, uint32_t NX, uint32_t NY, uint32_t NZ) {
    const uint32_t nxy = NX * NY;
    const uint32_t z = n / nxy;
    const uint32_t t = n - z * nxy;
    const uint32_t y = t / NX;
    const uint32_t x = t - y * NX;
    Nbr b;
    b.x0 = x;
    b.xp = (x + 1u == NX) ? 0u : x + 1u;
    b.xm = (x == 0u) ? NX - 1u : x - 1u;
    b.y0 = y * NX;
    b.yp = ((y + 1u == NY) ? 0u : y + 1u) * NX;
    b.ym = ((y == 0u) ? NY - 1u : y - 1u) * NX;
    b.z0 = z * nxy;
    b.zp = ((z + 1u == NZ) ? 0u : z + 1u) * nxy;
    b.zm = ((z == 0u) ? NZ - 1u : z - 1u) * nxy;
    return b;
}

template <int Q, int I> __device__ __forceinline__ uint32_t nb_index(const Nbr& b) {
    constexpr int cx = VS<Q>::cx[I], cy = VS<Q>::cy[I], cz = VS<Q>::cz[I];
    const uint32_t x = cx == 0 ? b.x0 : (cx > 0 ? b.xp : b.xm);
    const uint32_t y = cy == 0 ? b.y0 : (cy > 0 ? b.yp : b.ym);
    const uint32_t z = cz == 0 ? b.z0 : (cz > 0 ? b.zp : b.zm);
    return x + y + z;
}

// 4-vectors of the arithmetic type (LOD sums)
template <class R> struct V4;
template <> struct V4<float> { typedef float4 T; };
template <> struct V4<double> { typedef double4 T; };
template <class R> __device__ __forceinline__ typename V4<R>::T mk4(R x, R y, R z, R w) {
    typename V4<R>::T v;
    v.x = x;
    v.y = y;
    v.z = z;
    v.w = w;
    return v;
}

// Element idx of the array starting at `base` through a 32-bit byte offset,
// which gives the global load/store form with a scalar 64-bit base and a
// 32-bit vector offset (one VALU shift per distinct index instead of 64-bit
// address arithmetic per access). Valid while idx * sizeof(T) < 2^32: every
// array here has at most N elements per base, and the host rejects lattices
// with N > 2^30 (elements of at most 4 bytes) and, for the FP64
// instantiation (8-byte elements), N > 2^29.
template <class T> __device__ __forceinline__ T* at(T* base, uint32_t idx) {
    return (T*)((char*)base + (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <class T> __device__ __forceinline__ const T* at(const T* base, uint32_t idx) {
    return (const T*)((const char*)base + (uint32_t)(idx * (uint32_t)sizeof(T)));
}

// Esoteric-pull load/store of an NQ-direction lattice (load_f/store_f and
// load_a/store_a, sim_kernels.cl:234-247, 397-410). Direction pair (i, i+1)
// of cell n lives in slots (n, i+1) and (j[i], i) on even steps and in
// (n, i) and (j[i], i+1) on odd steps. Slot bases are uniform (odd is a
// kernel argument, 0 or 1, used arithmetically so that the compiler keeps
// the slot selection on the scalar unit); only the cell index varies per lane.
template <int Q, int NQ, class S>
__device__ __forceinline__ void load_ddfs(const typename S::T* __restrict__ buf, uint64_t N, uint32_t n,
                                          const Nbr& b, uint32_t odd, typename S::R* fh) {
    fh[0] = S::ld(*at(buf, n));
    static_for<(NQ - 1) / 2>([&](auto p) {
        constexpr int i = 2 * decltype(p)::value + 1;
        const uint32_t j = nb_index<Q, i>(b);
        const typename S::T* so = buf + (uint64_t)(i + 1u - odd) * N;
        const typename S::T* sn = buf + (uint64_t)(i + odd) * N;
        fh[i] = S::ld(*at(so, n));
        fh[i + 1] = S::ld(*at(sn, j));
    });
}
// Raw (undecoded) esoteric-pull loads, so that a kernel can issue the loads
// of all its lattices before the first decode and keep them in flight together.
template <int Q, int NQ, class S>
__device__ __forceinline__ void load_raw(const typename S::T* __restrict__ buf, uint64_t N, uint32_t n, const Nbr& b,
                                         uint32_t odd, typename S::T* raw) {
    raw[0] = *at(buf, n);
    static_for<(NQ - 1) / 2>([&](auto p) {
        constexpr int i = 2 * decltype(p)::value + 1;
        const uint32_t j = nb_index<Q, i>(b);
        const typename S::T* so = buf + (uint64_t)(i + 1u - odd) * N;
        const typename S::T* sn = buf + (uint64_t)(i + odd) * N;
        raw[i] = __builtin_nontemporal_load(at(so, n));
        raw[i + 1] = __builtin_nontemporal_load(at(sn, j));
    });
}
template <int NQ, class S>
__device__ __forceinline__ void decode(const typename S::T* raw, typename S::R* fh) {
#pragma unroll
    for (int i = 0; i < NQ; i++) fh[i] = S::ld(raw[i]);
}

template <int Q, int NQ, class S>
__device__ __forceinline__ void store_ddfs(typename S::T* __restrict__ buf, uint64_t N, uint32_t n, const Nbr& b,
                                           uint32_t odd, const typename S::R* fh) {
    *at(buf, n) = S::st(fh[0]);
    static_for<(NQ - 1) / 2>([&](auto p) {
        constexpr int i = 2 * decltype(p)::value + 1;
        const uint32_t j = nb_index<Q, i>(b);
        typename S::T* so = buf + (uint64_t)(i + 1u - odd) * N;
        typename S::T* sn = buf + (uint64_t)(i + odd) * N;
        __builtin_nontemporal_store(S::st(fh[i]), at(sn, j));
        __builtin_nontemporal_store(S::st(fh[i + 1]), at(so, n));
    });
}

// Stores whose slot lies one cell along +x (the directions i with c_x = +1
// store to the neighbour n + c_i): a wavefront's 64 lanes then write 64
// consecutive slots starting one element past a cache-line boundary. On this
// GPU the esoteric-pull pattern of 45 arrays at 256^3 runs at 5.2 TB/s with
// 4-byte elements and 5.5 TB/s with the same stores unshifted (2-byte: 4.5
// against 5.4 TB/s). Inside a block of whole rows
// (or of whole LOD segments of one row) every such store is handed to the
// work-item of the target cell through LDS instead, which writes it at its
// own, aligned position after one barrier; only a store that leaves the
// block (the last cell of a block that covers part of a row) is written
// directly. Each slot is still written once, with the same value.
constexpr uint32_t XCH_NONE = 0xffffffffu;
// position of odd direction I among the odd directions of the set with c_x > 0 (-1 if c_x == 0)
template <int Q, int I> __device__ __forceinline__ constexpr int xslot() {
    int k = 0;
    for (int i = 1; i < I; i += 2)
        if (VS<Q>::cx[i] > 0) k++;
    return VS<Q>::cx[I] > 0 ? k : -1;
}
template <int Q, int NQ> __device__ __forceinline__ constexpr int xslots() {
    int k = 0;
    for (int i = 1; i < NQ; i += 2)
        if (VS<Q>::cx[i] > 0) k++;
    return k;
}
// store_ddfs with the +x neighbour stores deposited in xch[(XB + slot) * W + dst]
// (dst: the target cell's work-item in the block, XCH_NONE: store directly)
template <int Q, int NQ, class S, int XB>
__device__ __forceinline__ void store_ddfs_x(typename S::T* __restrict__ buf, uint64_t N, uint32_t n, const Nbr& b,
                                             uint32_t odd, const typename S::R* fh, typename S::T* xch, uint32_t W,
                                             uint32_t dst) {
    *at(buf, n) = S::st(fh[0]);
    static_for<(NQ - 1) / 2>([&](auto p) {
        constexpr int i = 2 * decltype(p)::value + 1;
        const uint32_t j = nb_index<Q, i>(b);
        typename S::T* so = buf + (uint64_t)(i + 1u - odd) * N;
        typename S::T* sn = buf + (uint64_t)(i + odd) * N;
        constexpr int ks = xslot<Q, i>();
        if constexpr (ks >= 0) {
            if (dst != XCH_NONE) xch[(uint32_t)(XB + ks) * W + dst] = S::st(fh[i]);
            else __builtin_nontemporal_store(S::st(fh[i]), at(sn, j));
        } else {
            __builtin_nontemporal_store(S::st(fh[i]), at(sn, j));
        }
        __builtin_nontemporal_store(S::st(fh[i + 1]), at(so, n));
    });
}
// the stores handed to this cell: slot (i + odd, n + c_i - x) of every +x direction i,
// i.e. the neighbour index of the source cell, one cell back along x
template <int Q, int NQ, class S, int XB>
__device__ __forceinline__ void store_received(typename S::T* __restrict__ buf, uint64_t N, const Nbr& b,
                                               uint32_t odd, const typename S::T* xch, uint32_t W, uint32_t tid) {
    static_for<(NQ - 1) / 2>([&](auto p) {
        constexpr int i = 2 * decltype(p)::value + 1;
        constexpr int ks = xslot<Q, i>();
        if constexpr (ks >= 0) {
            constexpr int cy = VS<Q>::cy[i], cz = VS<Q>::cz[i];
            const uint32_t y = cy == 0 ? b.y0 : (cy > 0 ? b.yp : b.ym);
            const uint32_t z = cz == 0 ? b.z0 : (cz > 0 ? b.zp : b.zm);
            typename S::T* sn = buf + (uint64_t)(i + odd) * N;
            __builtin_nontemporal_store(xch[(uint32_t)(XB + ks) * W + tid], at(sn, b.x0 + y + z));
        }
    });
}

// 2-byte codes: store_ddfs with the +x neighbour stores' codes captured in
// cap[slot] instead of stored (they are packed in pairs into 32-bit LDS words
// once all three lattices are done), and the stores of captured codes
// (store_cap: the cell's own targets; store_received_codes: the targets of the
// cell one back along x, one cell back from this cell's neighbours)
template <int Q, int NQ, class S>
__device__ __forceinline__ void store_ddfs_cap(typename S::T* __restrict__ buf, uint64_t N, uint32_t n, const Nbr& b,
                                               uint32_t odd, const typename S::R* fh, typename S::T* cap) {
    *at(buf, n) = S::st(fh[0]);
    static_for<(NQ - 1) / 2>([&](auto p) {
        constexpr int i = 2 * decltype(p)::value + 1;
        const uint32_t j = nb_index<Q, i>(b);
        typename S::T* so = buf + (uint64_t)(i + 1u - odd) * N;
        typename S::T* sn = buf + (uint64_t)(i + odd) * N;
        constexpr int ks = xslot<Q, i>();
        if constexpr (ks >= 0) cap[ks] = S::st(fh[i]);
        else __builtin_nontemporal_store(S::st(fh[i]), at(sn, j));
        __builtin_nontemporal_store(S::st(fh[i + 1]), at(so, n));
    });
}
template <int Q, int NQ, class S>
__device__ __forceinline__ void store_cap(typename S::T* __restrict__ buf, uint64_t N, const Nbr& b, uint32_t odd,
                                          const typename S::T* cap) {
    static_for<(NQ - 1) / 2>([&](auto p) {
        constexpr int i = 2 * decltype(p)::value + 1;
        constexpr int ks = xslot<Q, i>();
        if constexpr (ks >= 0) {
            typename S::T* sn = buf + (uint64_t)(i + odd) * N;
            __builtin_nontemporal_store(cap[ks], at(sn, nb_index<Q, i>(b)));
        }
    });
}
template <int Q, int NQ, class S>
__device__ __forceinline__ void store_received_codes(typename S::T* __restrict__ buf, uint64_t N, const Nbr& b,
                                                     uint32_t odd, const typename S::T* code) {
    static_for<(NQ - 1) / 2>([&](auto p) {
        constexpr int i = 2 * decltype(p)::value + 1;
        constexpr int ks = xslot<Q, i>();
        if constexpr (ks >= 0) {
            constexpr int cy = VS<Q>::cy[i], cz = VS<Q>::cz[i];
            const uint32_t y = cy == 0 ? b.y0 : (cy > 0 ? b.yp : b.ym);
            const uint32_t z = cz == 0 ? b.z0 : (cz > 0 ? b.zp : b.zm);
            typename S::T* sn = buf + (uint64_t)(i + odd) * N;
            __builtin_nontemporal_store(code[ks], at(sn, b.x0 + y + z));
        }
    });
}

// ---------------------------------------------------------------------------
// initialize, sim_kernels.cl:760-832 (writes the t = 1 layout)
// ---------------------------------------------------------------------------
template <int Q, class S, bool MHD, bool ECR = false>
__global__ void __launch_bounds__(256) k_initialize(LbmArgs a) {
    typedef typename S::R R;
    typedef typename V4<R>::T V4;
    (void)sizeof(V4);
    R* const Rho = (R*)a.rho;
    R* const U = (R*)a.u;
    const R* const Fp = (const R*)a.F;
    R* const Ed = (R*)a.e_dyn;
    R* const Bd = (R*)a.b_dyn;
    const R* const Es = (const R*)a.e_stat;
    const R* const Bs = (const R*)a.b_stat;
    R* const Qp = (R*)a.q;
    V4* const Lseg = (V4*)a.lod_seg;
    const R W = R(a.w), C = R(a.c), KKGE = R(a.kkge), WQ = R(a.wq), ILS = sizeof(R) == 8 ? R(a.ils_d) : R(a.ils);
    const R FX = R(a.fx), FY = R(a.fy), FZ = R(a.fz);
    (void)Rho; (void)U; (void)Fp; (void)Ed; (void)Bd; (void)Es; (void)Bs; (void)Qp; (void)Lseg;
    (void)W; (void)C; (void)KKGE; (void)WQ; (void)ILS; (void)FX; (void)FY; (void)FZ;
    const uint64_t n64 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n64 >= a.N) return;
    const uint32_t n = (uint32_t)n64;
    const uint64_t N = a.N;
    const Nbr b = neighbours(n, a.nx, a.ny, a.nz);
    const uint8_t bo = a.flags[n] & 0x1F;
    if (bo == 0x01) {
        U[n] = R(0);
        U[N + n] = R(0);
        U[2 * N + n] = R(0);
        if (MHD) Qp[n] = R(0);
    }
    const R ux = U[n], uy = U[N + n], uz = U[2 * N + n];
    R fe[Q];
    f_eq<Q>(Rho[n], ux, uy, uz, fe);
    store_ddfs<Q, Q, S>((typename S::T*)a.fi, N, n, b, 1u, fe);
    if (MHD) {
        R qeq[7];
        a_eq(Qp[n], ux, uy, uz, qeq);
        store_ddfs<19, 7, S>((typename S::T*)a.fqi, N, n, b, 1u, qeq);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            Bd[k * N + n] = Bs[k * N + n];
            Ed[k * N + n] = Es[k * N + n];
        }
        // sim_kernels.cl:823-825 writes density 0; a non-zero ei_rho0 (build
        // extension) starts the electron gas of non-solid cells at that density
        f_eq<Q>(bo == 0x01 ? R(0) : R(a.ei_rho0), ux, uy, uz, fe);
        store_ddfs<Q, Q, S>((typename S::T*)a.ei, N, n, b, 1u, fe);
        if constexpr (ECR) {  // electron temperature DDFs, sim_kernels.cl:827-831
            a_eq(R(a.et[n]), ux, uy, uz, qeq);
            store_ddfs<19, 7, S>((typename S::T*)a.eti, N, n, b, 1u, qeq);
        }
    }
}

// ---------------------------------------------------------------------------
// stream_collide, sim_kernels.cl:465-758
// LOD: the block covers a contiguous range of whole LOD row segments
// (lod_bx consecutive cells of one row); the per-cell contributions are
// summed per segment by the pairwise tree in LDS and written to lod_seg.
// ---------------------------------------------------------------------------
// MHD variants with 32-bit arithmetic are held to 128 VGPRs (4 wavefronts per
// SIMD); with the neighbour offsets recomputed before the stores they need
// 111-114, and holding them to 96 (5 wavefronts) spills 14-20.
template <int Q, class S, bool MHD, bool TRT, bool LOD, bool ECR = false>
__global__ void __launch_bounds__(MHD ? 512 : 256)
    __attribute__((amdgpu_waves_per_eu(MHD && !ECR && sizeof(typename S::R) == 4 ? 4 : 1))) k_stream_collide(LbmArgs a) {
    typedef typename S::R R;
    typedef typename V4<R>::T V4;
    (void)sizeof(V4);
    R* const Rho = (R*)a.rho;
    R* const U = (R*)a.u;
    const R* const Fp = (const R*)a.F;
    R* const Ed = (R*)a.e_dyn;
    R* const Bd = (R*)a.b_dyn;
    const R* const Es = (const R*)a.e_stat;
    const R* const Bs = (const R*)a.b_stat;
    R* const Qp = (R*)a.q;
    V4* const Lseg = (V4*)a.lod_seg;
    const R W = R(a.w), C = R(a.c), KKGE = R(a.kkge), WQ = R(a.wq), ILS = sizeof(R) == 8 ? R(a.ils_d) : R(a.ils);
    const R FX = R(a.fx), FY = R(a.fy), FZ = R(a.fz);
    (void)Rho; (void)U; (void)Fp; (void)Ed; (void)Bd; (void)Es; (void)Bs; (void)Qp; (void)Lseg;
    (void)W; (void)C; (void)KKGE; (void)WQ; (void)ILS; (void)FX; (void)FY; (void)FZ;
    extern __shared__ __align__(32) unsigned char lds_raw[];
    V4* lds_lod = (V4*)lds_raw;
    const uint32_t tid = threadIdx.x;
    const uint64_t n64 = (uint64_t)blockIdx.x * blockDim.x + tid;
    const uint64_t N = a.N;
    const uint32_t n = (uint32_t)n64;
    bool active = n64 < N;
    uint8_t bo = 0x01;
    if (active) bo = *at(a.flags, n) & 0x1F;
    active = active && bo != 0x01;
    V4 lod = mk4<R>(R(0), R(0), R(0), R(0));
    typedef typename S::T T;
    // +x neighbour stores through LDS: the D3Q19 MHD variants with the fused LOD
    // tree (whose barrier they share). 4-byte storage (store_ddfs_x): the values
    // [NXS][W] are deposited as each lattice is stored. 2-byte codes: one 16-bit
    // LDS write per code measured 2 % slower than the straddling stores (FP16C
    // 256^3: 0.854 against 0.838 ms), so the 11 codes are held in registers
    // until the last lattice is stored and handed over as 6 words of two codes
    // ([6][W]; FP16C 0.833 -> 0.810 ms, FP16S 0.824 -> 0.755 ms, bit-identical).
    // Then one fluid flag per work-item.
    constexpr bool XCH = MHD && LOD && !ECR && Q == 19 && sizeof(T) == 4;
    constexpr bool XCH16 = MHD && LOD && !ECR && Q == 19 && sizeof(T) == 2;
    constexpr int NXS = XCH ? 2 * xslots<19, 19>() + xslots<19, 7>() : 1;
    const uint32_t BW = blockDim.x;
    T* const xch = (T*)(lds_raw + (size_t)BW * sizeof(V4));
    uint32_t* const xch32 = (uint32_t*)(lds_raw + (size_t)BW * sizeof(V4));  // [6][W] code pairs
    uint8_t* const xfl = XCH16 ? (uint8_t*)(xch32 + 6 * (size_t)BW) : (uint8_t*)(xch + (size_t)NXS * BW);
    T cap_f[5], cap_e[5], cap_q[1];  // XCH16: the +x stores' codes until the hand-over
    (void)xch32;
    (void)cap_f;
    (void)cap_e;
    (void)cap_q;
    // work-item of the cell one step along +x when it is in this block
    // (blocks hold whole rows, or whole LOD segments of one row)
    auto xdst = [&](uint32_t x) -> uint32_t {
        const uint32_t NX = a.nx;
        if (NX <= BW) return x + 1u == NX ? tid + 1u - NX : tid + 1u;
        return tid + 1u < BW ? tid + 1u : XCH_NONE;
    };
    (void)xdst;

    if (active) {
        const bool is_e = a.eqb && bo == 0x02;
        const Nbr b = neighbours(n, a.nx, a.ny, a.nz);
        // MHD: the neighbour offsets again before each store instead of nine
        // registers held from the loads to the last store (the 32-bit MHD
        // variants are at their 128-register limit); the empty asm keeps the
        // compiler from reusing b
        auto nb_store = [&]() -> Nbr {
            if constexpr (!MHD) return b;
            uint32_t m = n;
            __asm__ volatile("" : "+v"(m));
            return neighbours(m, a.nx, a.ny, a.nz);
        };
        // every load of the cell first: fi, ei, fqi, E_dyn, B_dyn
        T rf[Q], re[MHD ? Q : 1], rq[MHD ? 7 : 1], rt[ECR ? 7 : 1];
        R Bx = R(0), By = R(0), Bz = R(0), Ex = R(0), Ey = R(0), Ez = R(0);
        load_raw<Q, Q, S>((const T*)a.fi, N, n, b, a.odd, rf);
        if (MHD) {
            load_raw<Q, Q, S>((const T*)a.ei, N, n, b, a.odd, re);
            load_raw<19, 7, S>((const T*)a.fqi, N, n, b, a.odd, rq);
            if constexpr (ECR) load_raw<19, 7, S>((const T*)a.eti, N, n, b, a.odd, rt);
            Bx = *at(Bd, n); By = *at(Bd + N, n); Bz = *at(Bd + 2 * N, n);
            Ex = *at(Ed, n); Ey = *at(Ed + N, n); Ez = *at(Ed + 2 * N, n);
        }
        R fhn[Q];
        decode<Q, S>(rf, fhn);
        R rhon, uxn, uyn, uzn;
        if (is_e) {
            rhon = *at(Rho, n);
            uxn = *at(U, n);
            uyn = *at(U + N, n);
            uzn = *at(U + 2 * N, n);
        } else {
            rho_u<Q>(fhn, rhon, uxn, uyn, uzn);
        }
        R fxn = a.fx, fyn = a.fy, fzn = a.fz;
        R Fin[Q], feq[Q];
        const R w = W;
        const R c_tau = fmaR(w, R(-0.5), R(1));
        if (a.ffield) {
            fxn += *at(Fp, n);
            fyn += *at(Fp + N, n);
            fzn += *at(Fp + 2 * N, n);
        }
        if (MHD) {
            // electron gas part 1
            R ehn[Q];
            decode<Q, S>(re, ehn);
            R rhon_e, uxn_e, uyn_e, uzn_e;
            rho_u<Q>(ehn, rhon_e, uxn_e, uyn_e, uzn_e);
            // charge advection part 1 (D3Q7 subset of D3Q19 neighbours)
            R qhn[7];
            decode<7, S>(rq, qhn);
            R rhon_q = R(0);
#pragma unroll
            for (int i = 0; i < 7; i++) rhon_q += qhn[i];
            rhon_q += R(1);
            if constexpr (ECR) {
                // SUBGRID_ECR, sim_kernels.cl:557-629 (the cyclotron frequency and
                // relative absorption of :582-585 are computed there but unused)
                R ethn[7];
                decode<7, S>(rt, ethn);
                R Etn = R(0);
#pragma unroll
                for (int i = 0; i < 7; i++) Etn += ethn[i];
                Etn += R(1);
                const R Evx = *at(a.e_var, n), Evy = *at(a.e_var + N, n), Evz = *at(a.e_var + 2 * N, n);
                const R lb = sqrtf((Bx * Bx + By * By) + Bz * Bz);
                const R sp = ((Evx * Bx + Evy * By) + Evz * Bz) / (lb * lb);
                const R px = Evx - sp * Bx, py = Evy - sp * By, pz = Evz - sp * Bz;
                const R env = sqrtf((px * px + py * py) + pz * pz);
                Etn += R(a.keabs) / R(a.kkbme) * (rhon_e + R(0.00001f)) / KKGE * (env * env);
                // grad_mag_v(n, B_dyn), sim_kernels.cl:453-461: mag(c + e) - mag(c - e) / 2
                R g0 = R(0), g1 = R(0), g2 = R(0);
                const uint32_t xx = b.x0, yy = b.y0 / a.nx, zz = b.z0 / (a.nx * a.ny);
                if (!(xx == 0u || xx == a.nx - 1u || yy == 0u || yy == a.ny - 1u || zz == 0u || zz == a.nz - 1u)) {
                    auto mag = [&](uint32_t i) {
                        const R v0 = *at(Bd, i), v1 = *at(Bd + N, i), v2 = *at(Bd + 2 * N, i);
                        return sqrtf((v0 * v0 + v1 * v1) + v2 * v2);
                    };
                    const uint32_t sy = a.nx, sz = a.nx * a.ny;
                    g0 = mag(n + 1u) - mag(n - 1u) / R(2);
                    g1 = mag(n + sy) - mag(n - sy) / R(2);
                    g2 = mag(n + sz) - mag(n - sz) / R(2);
                }
                const R kt = R(a.kkbme) * Etn;
                const R dx = kt * g0 / lb, dy = kt * g1 / lb, dz = kt * g2 / lb;
                uxn_e += dx;
                uyn_e += dy;
                uzn_e += dz;
                Etn -= sqrtf((dx * dx + dy * dy) + dz * dz) / R(a.kkbme);
                R eteq[7];
                a_eq(Etn, uxn_e, uyn_e, uzn_e, eteq);
                if (a.upd) *at(a.et, n) = Etn;
                const R wq = WQ;
#pragma unroll
                for (int i = 0; i < 7; i++) ethn[i] = fmaR(R(1) - wq, ethn[i], wq * eteq[i]);
                store_ddfs<19, 7, S>((T*)a.eti, a.N_store, n, b, a.odd, ethn);
                const R dq = R(0.0001f) * Etn;  // ionisation
                rhon_e += dq;
                rhon_q += dq;
            }
            // charge inlet (build extension): a TYPE_E cell keeps the charge the
            // Q buffer holds and its charge DDFs are reset to that equilibrium
            const bool inlet = is_e && a.qin;
            if (inlet) rhon_q = *at(Qp, n);
            // charge advection part 2
            else *at(Qp, n) = rhon_q - rhon_e;
            R qeq[7];
            a_eq(rhon_q, uxn, uyn, uzn, qeq);
            const R wq = WQ;
#pragma unroll
            for (int i = 0; i < 7; i++) qhn[i] = inlet ? qeq[i] : fmaR(R(1) - wq, qhn[i], wq * qeq[i]);
            if constexpr (XCH) store_ddfs_x<19, 7, S, 2 * xslots<19, 19>()>((T*)a.fqi, a.N_store, n, nb_store(), a.odd, qhn, xch, BW, xdst(b.x0));
            else if constexpr (XCH16) store_ddfs_cap<19, 7, S>((T*)a.fqi, a.N_store, n, nb_store(), a.odd, qhn, cap_q);
            else store_ddfs<19, 7, S>((T*)a.fqi, a.N_store, n, nb_store(), a.odd, qhn);
            // electron gas part 2: F = -rho_e (E + u_e x B), Guo forcing, SRT
            const R crx = uyn_e * Bz - uzn_e * By;
            const R cry = uzn_e * Bx - uxn_e * Bz;
            const R crz = uxn_e * By - uyn_e * Bx;
            const R efx = -rhon_e * (Ex + crx), efy = -rhon_e * (Ey + cry), efz = -rhon_e * (Ez + crz);
            const R rho2_e = R(0.5) / (rhon_e * KKGE);
            uxn_e = clampR(fmaR(efx, rho2_e, uxn_e), -C, C);
            uyn_e = clampR(fmaR(efy, rho2_e, uyn_e), -C, C);
            uzn_e = clampR(fmaR(efz, rho2_e, uzn_e), -C, C);
            forcing<Q>(uxn_e, uyn_e, uzn_e, efx, efy, efz, Fin);
            f_eq<Q>(rhon_e, uxn_e, uyn_e, uzn_e, feq);
#pragma unroll
            for (int i = 0; i < Q; i++) {
                const R fi = Fin[i] * c_tau;
                ehn[i] = is_e ? feq[i] : fmaR(R(1) - w, ehn[i], fmaR(w, feq[i], fi));
            }
            if constexpr (XCH) store_ddfs_x<Q, Q, S, xslots<19, 19>()>((T*)a.ei, a.N_store, n, nb_store(), a.odd, ehn, xch, BW, xdst(b.x0));
            else if constexpr (XCH16) store_ddfs_cap<Q, Q, S>((T*)a.ei, a.N_store, n, nb_store(), a.odd, ehn, cap_e);
            else store_ddfs<Q, Q, S>((T*)a.ei, a.N_store, n, nb_store(), a.odd, ehn);
            // EM force on the gas (plain multiply-add, as written)
            fxn += rhon_q * (Ex + uyn * Bz - uzn * By);
            fyn += rhon_q * (Ey + uzn * Bx - uxn * Bz);
            fzn += rhon_q * (Ez + uxn * By - uyn * Bx);
            if (LOD) lod = mk4<R>(rhon_q - rhon_e, uxn * ILS, uyn * ILS, uzn * ILS);
        }
        if (a.vforce) {
            const R rho2 = R(0.5) / rhon;
            uxn = clampR(fmaR(fxn, rho2, uxn), -C, C);
            uyn = clampR(fmaR(fyn, rho2, uyn), -C, C);
            uzn = clampR(fmaR(fzn, rho2, uzn), -C, C);
            forcing<Q>(uxn, uyn, uzn, fxn, fyn, fzn, Fin);
        } else {
            uxn = clampR(uxn, -C, C);
            uyn = clampR(uyn, -C, C);
            uzn = clampR(uzn, -C, C);
#pragma unroll
            for (int i = 0; i < Q; i++) Fin[i] = R(0);
        }
        if (a.upd && !is_e) {
            *at(Rho, n) = rhon;
            *at(U, n) = uxn;
            *at(U + N, n) = uyn;
            *at(U + 2 * N, n) = uzn;
        }
        f_eq<Q>(rhon, uxn, uyn, uzn, feq);
        if (!TRT) {
            const bool vf = a.vforce != 0;
#pragma unroll
            for (int i = 0; i < Q; i++) {
                const R fi = vf ? Fin[i] * c_tau : Fin[i];
                fhn[i] = is_e ? feq[i] : fmaR(R(1) - w, fhn[i], fmaR(w, feq[i], fi));
            }
        } else {
            const R wp = w;
            const R wm = R(1) / (R(0.1875) / (R(1) / w - R(0.5)) + R(0.5));
            if (a.vforce) {
                const R c_taup = fmaR(wp, R(-0.25), R(0.5)), c_taum = fmaR(wm, R(-0.25), R(0.5));
                R Fib[Q];
                Fib[0] = Fin[0];
#pragma unroll
                for (int i = 1; i < Q; i += 2) {
                    Fib[i] = Fin[i + 1];
                    Fib[i + 1] = Fin[i];
                }
#pragma unroll
                for (int i = 0; i < Q; i++) Fin[i] = fmaR(c_taup, Fin[i] + Fib[i], c_taum * (Fin[i] - Fib[i]));
            }
            R fhb[Q], feb[Q];
            fhb[0] = fhn[0];
            feb[0] = feq[0];
#pragma unroll
            for (int i = 1; i < Q; i += 2) {
                fhb[i] = fhn[i + 1];
                fhb[i + 1] = fhn[i];
                feb[i] = feq[i + 1];
                feb[i + 1] = feq[i];
            }
#pragma unroll
            for (int i = 0; i < Q; i++)
                fhn[i] = is_e ? feq[i]
                              : fmaR(R(0.5) * wp, feq[i] - fhn[i] + feb[i] - fhb[i],
                                               fmaR(R(0.5) * wm, feq[i] - feb[i] - fhn[i] + fhb[i], fhn[i] + Fin[i]));
        }
        if constexpr (XCH) store_ddfs_x<Q, Q, S, 0>((T*)a.fi, a.N_store, n, nb_store(), a.odd, fhn, xch, BW, xdst(b.x0));
        else if constexpr (XCH16) {
            store_ddfs_cap<Q, Q, S>((T*)a.fi, a.N_store, n, nb_store(), a.odd, fhn, cap_f);
            const uint32_t dst = xdst(b.x0);
            if (dst != XCH_NONE) {  // to the work-item of cell n + 1 along x, two codes per word
                auto pk = [](uint16_t lo, uint16_t hi) { return (uint32_t)lo | ((uint32_t)hi << 16); };
                xch32[0 * BW + dst] = pk(cap_f[0], cap_f[1]);
                xch32[1 * BW + dst] = pk(cap_f[2], cap_f[3]);
                xch32[2 * BW + dst] = pk(cap_f[4], cap_e[0]);
                xch32[3 * BW + dst] = pk(cap_e[1], cap_e[2]);
                xch32[4 * BW + dst] = pk(cap_e[3], cap_e[4]);
                xch32[5 * BW + dst] = (uint32_t)cap_q[0];
            } else {  // the target leaves the block: store directly
                const Nbr bs = nb_store();
                store_cap<Q, Q, S>((T*)a.fi, a.N_store, bs, a.odd, cap_f);
                store_cap<Q, Q, S>((T*)a.ei, a.N_store, bs, a.odd, cap_e);
                store_cap<19, 7, S>((T*)a.fqi, a.N_store, bs, a.odd, cap_q);
            }
        }
        else store_ddfs<Q, Q, S>((T*)a.fi, a.N_store, n, nb_store(), a.odd, fhn);
    }

    if (LOD) {
        // level-1 pairwise tree over each LOD row segment
        const uint32_t bx = a.lod_bx;
        lds_lod[tid] = lod;
        if constexpr (XCH || XCH16) xfl[tid] = active ? 1 : 0;
        __syncthreads();
        if constexpr (XCH16) {
            if (n64 < N) {
                const Nbr br = neighbours(n, a.nx, a.ny, a.nz);
                const uint32_t NX = a.nx;
                const uint32_t src = NX <= BW ? (br.x0 == 0u ? tid + NX - 1u : tid - 1u) : (tid > 0u ? tid - 1u : XCH_NONE);
                if (src != XCH_NONE && xfl[src]) {
                    uint32_t w[6];
#pragma unroll
                    for (int k = 0; k < 6; k++) w[k] = xch32[k * BW + tid];
                    const T rf[5] = {(T)w[0], (T)(w[0] >> 16), (T)w[1], (T)(w[1] >> 16), (T)w[2]};
                    const T re[5] = {(T)(w[2] >> 16), (T)w[3], (T)(w[3] >> 16), (T)w[4], (T)(w[4] >> 16)};
                    const T rq[1] = {(T)w[5]};
                    store_received_codes<Q, Q, S>((T*)a.fi, a.N_store, br, a.odd, rf);
                    store_received_codes<Q, Q, S>((T*)a.ei, a.N_store, br, a.odd, re);
                    store_received_codes<19, 7, S>((T*)a.fqi, a.N_store, br, a.odd, rq);
                }
            }
        }
        if constexpr (XCH) {
            // the +x neighbour stores deposited for this cell by the cell one back
            // along x (if it is in the block and not solid)
            if (n64 < N) {
                const Nbr br = neighbours(n, a.nx, a.ny, a.nz);
                const uint32_t NX = a.nx;
                const uint32_t src = NX <= BW ? (br.x0 == 0u ? tid + NX - 1u : tid - 1u) : (tid > 0u ? tid - 1u : XCH_NONE);
                if (src != XCH_NONE && xfl[src]) {
                    store_received<Q, Q, S, 0>((T*)a.fi, a.N_store, br, a.odd, xch, BW, tid);
                    store_received<Q, Q, S, xslots<19, 19>()>((T*)a.ei, a.N_store, br, a.odd, xch, BW, tid);
                    store_received<19, 7, S, 2 * xslots<19, 19>()>((T*)a.fqi, a.N_store, br, a.odd, xch, BW, tid);
                }
            }
        }
        const uint32_t i = tid % bx;
        if (bx <= 64u && (bx & (bx - 1u)) == 0u) {
            // segments of a power of two up to 64 cells lie inside one
            // wavefront: the same pairwise tree (same pairs, same order, so
            // the same sums) with cross-lane reads instead of LDS and four
            // barriers (FP16C 256^3: 0.794 -> 0.780 ms, FP16S 0.758 -> 0.751)
            V4 v = lod;
            for (uint32_t s = bx >> 1; s >= 1; s >>= 1) {
                const V4 o = mk4<R>(__shfl_down(v.x, s), __shfl_down(v.y, s), __shfl_down(v.z, s), __shfl_down(v.w, s));
                if (i < s) v = mk4<R>(v.x + o.x, v.y + o.y, v.z + o.z, v.w + o.w);
            }
            if (i == 0 && n64 < N) Lseg[n64 / bx] = v;
        } else {
            uint32_t P = 1;
            while (P < bx) P <<= 1;
            for (uint32_t s = P >> 1; s >= 1; s >>= 1) {
                if (i < s) {
                    const V4 v = lds_lod[tid];
                    const V4 o = (i + s < bx) ? lds_lod[tid + s] : mk4<R>(R(0), R(0), R(0), R(0));
                    lds_lod[tid] = mk4<R>(v.x + o.x, v.y + o.y, v.z + o.z, v.w + o.w);
                }
                __syncthreads();
            }
            if (i == 0 && n64 < N) Lseg[n64 / bx] = lds_lod[tid];
        }

    }
}

// ---------------------------------------------------------------------------
// update_fields, sim_kernels.cl:834-859
// ---------------------------------------------------------------------------
template <int Q, class S>
__global__ void __launch_bounds__(256) k_update_fields(LbmArgs a) {
    typedef typename S::R R;
    typedef typename V4<R>::T V4;
    (void)sizeof(V4);
    R* const Rho = (R*)a.rho;
    R* const U = (R*)a.u;
    const R* const Fp = (const R*)a.F;
    R* const Ed = (R*)a.e_dyn;
    R* const Bd = (R*)a.b_dyn;
    const R* const Es = (const R*)a.e_stat;
    const R* const Bs = (const R*)a.b_stat;
    R* const Qp = (R*)a.q;
    V4* const Lseg = (V4*)a.lod_seg;
    const R W = R(a.w), C = R(a.c), KKGE = R(a.kkge), WQ = R(a.wq), ILS = sizeof(R) == 8 ? R(a.ils_d) : R(a.ils);
    const R FX = R(a.fx), FY = R(a.fy), FZ = R(a.fz);
    (void)Rho; (void)U; (void)Fp; (void)Ed; (void)Bd; (void)Es; (void)Bs; (void)Qp; (void)Lseg;
    (void)W; (void)C; (void)KKGE; (void)WQ; (void)ILS; (void)FX; (void)FY; (void)FZ;
    const uint64_t n64 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n64 >= a.N) return;
    const uint32_t n = (uint32_t)n64;
    const uint64_t N = a.N;
    if ((a.flags[n] & 0x1F) == 0x01) return;
    const Nbr b = neighbours(n, a.nx, a.ny, a.nz);
    R fhn[Q];
    load_ddfs<Q, Q, S>((const typename S::T*)a.fi, N, n, b, a.odd, fhn);
    R rhon, uxn, uyn, uzn;
    rho_u<Q>(fhn, rhon, uxn, uyn, uzn);
    Rho[n] = rhon;
    U[n] = clampR(uxn, -C, C);
    U[N + n] = clampR(uyn, -C, C);
    U[2 * N + n] = clampR(uzn, -C, C);
}

// ---------------------------------------------------------------------------
// Level-2 LOD tree: block L sums the by*bz row-segment sums of LOD L
// (index j = yy + zz*by) by the pairwise tree, and writes QU_lod[L]
// (top-level LODs at offset 0 for a single domain, sim_kernels.cl:667-671).
// ---------------------------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(1024) k_lod_reduce(const typename V4<R>::T* __restrict__ seg,
                                                     typename V4<R>::T* __restrict__ qu, uint32_t NY, uint32_t nd,
                                                     uint32_t by, uint32_t bz, uint32_t P) {
    typedef typename V4<R>::T V;
    extern __shared__ __align__(32) unsigned char lds_raw[];
    V* a4 = (V*)lds_raw;
    const uint32_t L = blockIdx.x;
    const uint32_t X = L % nd, Y = (L / nd) % nd, Z = L / (nd * nd);
    const uint32_t cnt = by * bz;
    for (uint32_t j = threadIdx.x; j < P; j += blockDim.x) {
        V v = mk4<R>(R(0), R(0), R(0), R(0));
        if (j < cnt) {
            const uint32_t yy = j % by, zz = j / by;
            const uint64_t row = (uint64_t)(Z * bz + zz) * NY + (Y * by + yy);
            v = seg[row * nd + X];
        }
        a4[j] = v;
    }
    __syncthreads();
    for (uint32_t s = P >> 1; s >= 1; s >>= 1) {
        for (uint32_t i = threadIdx.x; i < s; i += blockDim.x) {
            const V v = a4[i], o = a4[i + s];
            a4[i] = mk4<R>(v.x + o.x, v.y + o.y, v.z + o.z, v.w + o.w);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) qu[L] = a4[0];
}

template <class R> __global__ void k_fill(R* p, uint64_t n, R v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Device codecs on arrays (diagnostics: lets tests compare the device
// conversion of every value class with the oracle's).
template <class S>
__global__ void k_codec(int dir, const void* in, void* out, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dir == 0) ((uint16_t*)out)[i] = S::st(((const float*)in)[i]);
    else ((float*)out)[i] = S::ld(((const uint16_t*)in)[i]);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

template <int Q, class S>
static hipError_t init_impl(bool mhd, const LbmArgs& a, hipStream_t s) {
    const uint32_t g = blocks_for(a.N, 256);
    if (mhd) hipLaunchKernelGGL((k_initialize<Q, S, true>), dim3(g), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_initialize<Q, S, false>), dim3(g), dim3(256), 0, s, a);
    return hipGetLastError();
}

// SUBGRID_ECR variants (D3Q19, MHD, 32-bit arithmetic)
template <class S>
static hipError_t init_ecr_impl(const LbmArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((k_initialize<19, S, true, true>), dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}
template <class S>
static hipError_t sc_ecr_impl(bool trt, bool lod, uint32_t block, const LbmArgs& a, hipStream_t s) {
    const uint32_t g = blocks_for(a.N, block);
    const size_t sh = lod ? (size_t)block * sizeof(float4) : 0;
#define ION_SC(T, L) hipLaunchKernelGGL((k_stream_collide<19, S, true, T, L, true>), dim3(g), dim3(block), sh, s, a)
    if (lod) { if (trt) ION_SC(true, true); else ION_SC(false, true); }
    else { if (trt) ION_SC(true, false); else ION_SC(false, false); }
#undef ION_SC
    return hipGetLastError();
}

template <int Q, class S>
static hipError_t sc_impl(bool mhd, bool trt, bool lod, uint32_t block, const LbmArgs& a, hipStream_t s) {
    const uint32_t g = blocks_for(a.N, block);
    // LOD partials, and for the D3Q19 MHD variants the +x stores handed over through LDS
    // (11 values: 5 of fi, 5 of ei, 1 of fqi) plus one fluid flag per work-item
    constexpr size_t xch = Q != 19 ? 0 : (sizeof(typename S::T) == 4 ? 11 * 4 + 1 : 6 * 4 + 1);
    const size_t sh = lod ? (size_t)block * (sizeof(typename V4<typename S::R>::T) + (mhd ? xch : 0)) : 0;
#define ION_SC(M, T, L) hipLaunchKernelGGL((k_stream_collide<Q, S, M, T, L>), dim3(g), dim3(block), sh, s, a)
    if (mhd) {
        if (lod) { if (trt) ION_SC(true, true, true); else ION_SC(true, false, true); }
        else { if (trt) ION_SC(true, true, false); else ION_SC(true, false, false); }
    } else {
        if (trt) ION_SC(false, true, false); else ION_SC(false, false, false);
    }
#undef ION_SC
    return hipGetLastError();
}

template <int Q, class S>
static hipError_t uf_impl(const LbmArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((k_update_fields<Q, S>), dim3(blocks_for(a.N, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_initialize(int q, int store, bool mhd, const LbmArgs& a, hipStream_t s, bool ecr) {
    if (ecr) {
        if (q != 19 || !mhd) return hipErrorInvalidValue;
        if (store == STORE_FP32) return init_ecr_impl<F32>(a, s);
        if (store == STORE_FP16S) return init_ecr_impl<F16S>(a, s);
        if (store == STORE_FP16C) return init_ecr_impl<F16C>(a, s);
        return hipErrorInvalidValue;
    }
    if (store == STORE_FP64) return q == 19 ? init_impl<19, F64>(mhd, a, s) : hipErrorInvalidValue;
    if (q == 19) {
        if (store == STORE_FP32) return init_impl<19, F32>(mhd, a, s);
        if (store == STORE_FP16S) return init_impl<19, F16S>(mhd, a, s);
        if (store == STORE_FP16C) return init_impl<19, F16C>(mhd, a, s);
    } else if (q == 15) {
        if (store == STORE_FP32) return init_impl<15, F32>(mhd, a, s);
        if (store == STORE_FP16S) return init_impl<15, F16S>(mhd, a, s);
        if (store == STORE_FP16C) return init_impl<15, F16C>(mhd, a, s);
    } else if (q == 9 && !mhd) {
        if (store == STORE_FP32) return init_impl<9, F32>(false, a, s);
        if (store == STORE_FP16S) return init_impl<9, F16S>(false, a, s);
        if (store == STORE_FP16C) return init_impl<9, F16C>(false, a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_stream_collide(int q, int store, bool mhd, bool trt, bool lod, uint32_t block, const LbmArgs& a,
                                 hipStream_t s, bool ecr) {
    if (ecr) {
        if (q != 19 || !mhd) return hipErrorInvalidValue;
        if (store == STORE_FP32) return sc_ecr_impl<F32>(trt, lod, block, a, s);
        if (store == STORE_FP16S) return sc_ecr_impl<F16S>(trt, lod, block, a, s);
        if (store == STORE_FP16C) return sc_ecr_impl<F16C>(trt, lod, block, a, s);
        return hipErrorInvalidValue;
    }
    if (store == STORE_FP64) return q == 19 ? sc_impl<19, F64>(mhd, trt, lod, block, a, s) : hipErrorInvalidValue;
    if (q == 19) {
        if (store == STORE_FP32) return sc_impl<19, F32>(mhd, trt, lod, block, a, s);
        if (store == STORE_FP16S) return sc_impl<19, F16S>(mhd, trt, lod, block, a, s);
        if (store == STORE_FP16C) return sc_impl<19, F16C>(mhd, trt, lod, block, a, s);
    } else if (q == 15) {
        if (store == STORE_FP32) return sc_impl<15, F32>(mhd, trt, lod, block, a, s);
        if (store == STORE_FP16S) return sc_impl<15, F16S>(mhd, trt, lod, block, a, s);
        if (store == STORE_FP16C) return sc_impl<15, F16C>(mhd, trt, lod, block, a, s);
    } else if (q == 9 && !mhd) {
        if (store == STORE_FP32) return sc_impl<9, F32>(false, trt, false, block, a, s);
        if (store == STORE_FP16S) return sc_impl<9, F16S>(false, trt, false, block, a, s);
        if (store == STORE_FP16C) return sc_impl<9, F16C>(false, trt, false, block, a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_update_fields(int q, int store, const LbmArgs& a, hipStream_t s) {
    if (store == STORE_FP64) return q == 19 ? uf_impl<19, F64>(a, s) : hipErrorInvalidValue;
    if (q == 19) {
        if (store == STORE_FP32) return uf_impl<19, F32>(a, s);
        if (store == STORE_FP16S) return uf_impl<19, F16S>(a, s);
        if (store == STORE_FP16C) return uf_impl<19, F16C>(a, s);
    } else if (q == 15) {
        if (store == STORE_FP32) return uf_impl<15, F32>(a, s);
        if (store == STORE_FP16S) return uf_impl<15, F16S>(a, s);
        if (store == STORE_FP16C) return uf_impl<15, F16C>(a, s);
    } else if (q == 9) {
        if (store == STORE_FP32) return uf_impl<9, F32>(a, s);
        if (store == STORE_FP16S) return uf_impl<9, F16S>(a, s);
        if (store == STORE_FP16C) return uf_impl<9, F16C>(a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_lod_reduce(const void* seg, void* qu_lod, bool f64, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t D,
                             hipStream_t s) {
    const uint32_t nd = 1u << D;
    const uint32_t by = ny / nd, bz = nz / nd;
    (void)nx;
    (void)nz;
    uint32_t P = 1;
    while (P < by * bz) P <<= 1;
    const size_t vs = f64 ? sizeof(double4) : sizeof(float4);
    if ((size_t)P * vs > 64 * 1024) return hipErrorInvalidValue;
    const uint32_t threads = P < 1024 ? (P < 64 ? 64 : P) : 1024;
    if (f64)
        hipLaunchKernelGGL(k_lod_reduce<double>, dim3(nd * nd * nd), dim3(threads), (size_t)P * vs, s,
                           (const double4*)seg, (double4*)qu_lod, ny, nd, by, bz, P);
    else
        hipLaunchKernelGGL(k_lod_reduce<float>, dim3(nd * nd * nd), dim3(threads), (size_t)P * vs, s,
                           (const float4*)seg, (float4*)qu_lod, ny, nd, by, bz, P);
    return hipGetLastError();
}

hipError_t launch_fill(void* p, bool f64, uint64_t n, double v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (f64) hipLaunchKernelGGL(k_fill<double>, dim3(blocks_for(n, 256)), dim3(256), 0, s, (double*)p, n, v);
    else hipLaunchKernelGGL(k_fill<float>, dim3(blocks_for(n, 256)), dim3(256), 0, s, (float*)p, n, (float)v);
    return hipGetLastError();
}

hipError_t launch_codec(int store, int dir, const void* in, void* out, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 g(blocks_for(n, 256)), b(256);
    if (store == STORE_FP16S) hipLaunchKernelGGL(k_codec<F16S>, g, b, 0, s, dir, in, out, n);
    else if (store == STORE_FP16C) hipLaunchKernelGGL(k_codec<F16C>, g, b, 0, s, dir, in, out, n);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace ion
