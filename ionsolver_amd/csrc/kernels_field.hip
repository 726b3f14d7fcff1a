// gfx950 kernels of the dynamic E/B field update (update_e_b_dynamic,
// sim_kernels.cl:897-993) and of the float64 conservation sums.
//
// The field at every fluid cell is a sum over the level-of-detail (LOD)
// sources of the own domain: E += q r/|r|^3, B += q v x r/|r|^3. The
// sources are compacted into a list (centre, charge, mean velocity, LOD
// index) once per step, then staged through LDS in tiles; every work-item
// accumulates the fields of its cells in registers, sweeping the sources
// in ascending LOD index as the reference loop does.
#include <hip/hip_runtime.h>

#include "cr_math.hpp"
#include "kernels.hpp"

namespace ion {

// ---------------------------------------------------------------------------
// Source list: LOD indices d in [lo, hi) with a non-zero charge, in index
// order. A source of zero charge adds exactly zero to every cell, so
// dropping it leaves all fields unchanged.
// ---------------------------------------------------------------------------
// 4-vectors of the arithmetic type; the 4th component of a source's vd
// carries its LOD index (float: the bits, double: the value, exact below 2^53).
template <class R> struct V4F;
template <> struct V4F<float> {
    typedef float4 T;
    static __device__ __forceinline__ float idx(uint32_t d) { return __uint_as_float(d); }
    static __device__ __forceinline__ uint32_t idx_of(float v) { return __float_as_uint(v); }
};
template <> struct V4F<double> {
    typedef double4 T;
    static __device__ __forceinline__ double idx(uint32_t d) { return (double)d; }
    static __device__ __forceinline__ uint32_t idx_of(double v) { return (uint32_t)v; }
};
template <class R> __device__ __forceinline__ typename V4F<R>::T mk4f(R x, R y, R z, R w) {
    typename V4F<R>::T v;
    v.x = x;
    v.y = y;
    v.z = z;
    v.w = w;
    return v;
}

template <class R>
__global__ void __launch_bounds__(1024) k_prep_sources(const typename V4F<R>::T* __restrict__ qu, uint32_t lo,
                                                       uint32_t hi, uint32_t nd, R dsx, R dsy, R dsz,
                                                       typename V4F<R>::T* __restrict__ src_pq,
                                                       typename V4F<R>::T* __restrict__ src_vd,
                                                       uint32_t* __restrict__ count) {
    __shared__ uint32_t wave_tot[16];
    __shared__ uint32_t base_sh;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
    if (tid == 0) base_sh = 0;
    __syncthreads();
    for (uint32_t c0 = lo; c0 < hi; c0 += 1024) {
        const uint32_t d = c0 + tid;
        typename V4F<R>::T e = mk4f<R>(R(0), R(0), R(0), R(0));
        bool keep = false;
        if (d < hi) {
            e = qu[d];
            keep = e.x != R(0);
        }
        const uint64_t mask = __ballot(keep);
        const uint32_t pre = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) wave_tot[wv] = __popcll(mask);
        __syncthreads();
        uint32_t off = base_sh;
        for (uint32_t k = 0; k < wv; k++) off += wave_tot[k];
        if (keep) {
            const uint32_t o = off + pre;
            const uint32_t t = d % (nd * nd);
            // lod_coordinates, sim_kernels.cl:442-447
            const R lx = R((float)(t % nd)) * dsx + R(0.5) * dsx;
            const R ly = R((float)(t / nd)) * dsy + R(0.5) * dsy;
            const R lz = R((float)(d / (nd * nd))) * dsz + R(0.5) * dsz;
            src_pq[o] = mk4f<R>(lx, ly, lz, e.x);
            src_vd[o] = mk4f<R>(e.y, e.z, e.w, V4F<R>::idx(d));
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t tot = 0;
            for (uint32_t k = 0; k < 16; k++) tot += wave_tot[k];
            base_sh += tot;
        }
        __syncthreads();
    }
    if (tid == 0) *count = base_sh;
}

constexpr uint32_t FIELD_TILE = 512;
constexpr uint32_t FIELD_BLOCK = 256;

// One interaction in the reference's arithmetic: pre = r / (sqrt(r.r))^3
// with correctly rounded sqrt and division, e += q*pre, b += q*(v x pre).
template <class R>
__device__ __forceinline__ void interact_exact(R q, R vx, R vy, R vz, R rx, R ry, R rz, R& ex, R& ey, R& ez, R& bx,
                                               R& by, R& bz) {
    const R s = sqrt((rx * rx + ry * ry) + rz * rz);
    const R cbm = (s * s) * s;
    const R prx = rx / cbm, pry = ry / cbm, prz = rz / cbm;
    ex = ex + q * prx;
    ey = ey + q * pry;
    ez = ez + q * prz;
    const R crx = vy * prz - vz * pry, cry = vz * prx - vx * prz, crz = vx * pry - vy * prx;
    bx = bx + q * crx;
    by = by + q * cry;
    bz = bz + q * crz;
}

// Judge partials of a block: the work-items' four sums (E error, E norm,
// B error, B norm) reduced in a fixed order (wavefront tree, then the
// wavefronts in order) into partial[blockIdx.x][4], in double.
__device__ __forceinline__ void judge_block_sums(const float v[4], double* partial) {
    __shared__ double jsh[16][4];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    double d[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        d[k] = (double)v[k];
        for (int o = 32; o >= 1; o >>= 1) d[k] += __shfl_down(d[k], o, 64);
    }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 4; k++) jsh[wv][k] = d[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (uint32_t w = 0; w < (blockDim.x + 63u) / 64u; w++) t += jsh[w][threadIdx.x];
        partial[(size_t)blockIdx.x * 4 + threadIdx.x] = t;
    }
}

template <bool FAST, bool NEAR, int CPT>
__global__ void __launch_bounds__(FIELD_BLOCK) k_update_e_b(FieldArgs a) {
    __shared__ float4 s_pq[FIELD_TILE];
    __shared__ float4 s_vd[FIELD_TILE];
    const uint32_t tid = threadIdx.x;
    const uint32_t NX = a.nx, NY = a.ny, NZ = a.nz;
    const uint32_t nd = 1u << a.D;
    const uint32_t lbx = NX / nd, lby = NY / nd, lbz = NZ / nd;

    float px[CPT], py[CPT], pz[CPT];
    float ex[CPT], ey[CPT], ez[CPT], bx[CPT], by[CPT], bz[CPT];
    uint32_t own[CPT], cell[CPT];
    bool act[CPT];
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        const uint64_t n64 = (uint64_t)blockIdx.x * (FIELD_BLOCK * CPT) + (uint64_t)k * FIELD_BLOCK + tid;
        act[k] = n64 < a.N;
        const uint32_t n = act[k] ? (uint32_t)n64 : 0u;
        cell[k] = n;
        if (act[k]) act[k] = (a.flags[n] & 0x1F) != 0x01;
        const uint32_t z = n / (NX * NY), t = n - z * NX * NY, y = t / NX, x = t - y * NX;
        px[k] = (float)x;
        py[k] = (float)y;
        pz[k] = (float)z;
        own[k] = x / lbx + (y / lby + z / lbz * nd) * nd;  // lod_index, sim_kernels.cl:425-433
        ex[k] = ey[k] = ez[k] = bx[k] = by[k] = bz[k] = 0.0f;
        if (NEAR && act[k]) {
            // individual cells of the own near-field block, x, y, z nesting
            // as in sim_kernels.cl:913-938
            const uint32_t x0 = (x / a.dsx) * a.dsx, y0 = (y / a.dsy) * a.dsy, z0 = (z / a.dsz) * a.dsz;
            const uint32_t xu = min(x0 + a.dsx, NX), yu = min(y0 + a.dsy, NY), zu = min(z0 + a.dsz, NZ);
            for (uint32_t xx = x0; xx < xu; xx++)
                for (uint32_t yy = y0; yy < yu; yy++)
                    for (uint32_t zz = z0; zz < zu; zz++) {
                        const uint32_t nc = xx + (yy + zz * NY) * NX;
                        if (nc == n) continue;
                        const float qc = a.q[nc];
                        if (qc == 0.0f) continue;
                        interact_exact(qc, a.u[nc], a.u[(uint64_t)nc + a.N], a.u[(uint64_t)nc + 2 * a.N],
                                       px[k] - (float)xx, py[k] - (float)yy, pz[k] - (float)zz, ex[k], ey[k], ez[k],
                                       bx[k], by[k], bz[k]);
                    }
        }
    }

    const uint32_t count = *a.src_count;
    for (uint32_t base = 0; base < count; base += FIELD_TILE) {
        const uint32_t m = min(FIELD_TILE, count - base);
        __syncthreads();
        for (uint32_t i = tid; i < m; i += FIELD_BLOCK) {
            s_pq[i] = a.src_pq[base + i];
            s_vd[i] = a.src_vd[base + i];
        }
        __syncthreads();
        for (uint32_t s = 0; s < m; s++) {
            const float4 pq = s_pq[s];
            const float4 vd = s_vd[s];
            const uint32_t d = __float_as_uint(vd.w);
#pragma unroll
            for (int k = 0; k < CPT; k++) {
                const float rx = px[k] - pq.x, ry = py[k] - pq.y, rz = pz[k] - pq.z;
                if constexpr (FAST) {
                    const float r2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, rz * rz));
                    const float inv = __builtin_amdgcn_rsqf(r2);
                    float t = pq.w * (inv * inv * inv);
                    t = (d == own[k]) ? 0.0f : t;
                    ex[k] = __builtin_fmaf(t, rx, ex[k]);
                    ey[k] = __builtin_fmaf(t, ry, ey[k]);
                    ez[k] = __builtin_fmaf(t, rz, ez[k]);
                    const float crx = __builtin_fmaf(vd.y, rz, -vd.z * ry);
                    const float cry = __builtin_fmaf(vd.z, rx, -vd.x * rz);
                    const float crz = __builtin_fmaf(vd.x, ry, -vd.y * rx);
                    bx[k] = __builtin_fmaf(t, crx, bx[k]);
                    by[k] = __builtin_fmaf(t, cry, by[k]);
                    bz[k] = __builtin_fmaf(t, crz, bz[k]);
                } else {
                    if (d != own[k])
                        interact_exact(pq.w, vd.x, vd.y, vd.z, rx, ry, rz, ex[k], ey[k], ez[k], bx[k], by[k], bz[k]);
                }
            }
        }
    }

    const uint64_t N = a.N;
    float jd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        if (!act[k]) continue;
        const uint32_t n = cell[k];
        if (a.raw != nullptr) {
            float* r = a.raw;
            r[n] = ex[k];
            r[N + n] = ey[k];
            r[2 * N + n] = ez[k];
            r[3 * N + n] = bx[k];
            r[4 * N + n] = by[k];
            r[5 * N + n] = bz[k];
            continue;
        }
        if (a.judge != nullptr) {
            const float d[6] = {ex[k], ey[k], ez[k], bx[k], by[k], bz[k]};
#pragma unroll
            for (int c = 0; c < 6; c++) {
                const float sur = (c < 3 ? a.e_dyn : a.b_dyn)[(uint64_t)(c % 3) * N + n];
                jd[c < 3 ? 0 : 2] = __builtin_fmaf(sur - d[c], sur - d[c], jd[c < 3 ? 0 : 2]);
                jd[c < 3 ? 1 : 3] = __builtin_fmaf(d[c], d[c], jd[c < 3 ? 1 : 3]);
            }
        }
        a.e_dyn[n] = a.e_stat[n] + a.ke * ex[k];
        a.e_dyn[N + n] = a.e_stat[N + n] + a.ke * ey[k];
        a.e_dyn[2 * N + n] = a.e_stat[2 * N + n] + a.ke * ez[k];
        a.b_dyn[n] = a.b_stat[n] + a.kmu * bx[k];
        a.b_dyn[N + n] = a.b_stat[N + n] + a.kmu * by[k];
        a.b_dyn[2 * N + n] = a.b_stat[2 * N + n] + a.kmu * bz[k];
    }
    if (a.judge != nullptr) judge_block_sums(jd, a.judge);
}

// FP64 instantiation of the field update (the error reference of the
// storage formats): one cell per work-item, the reference's arithmetic in
// double, sources staged through LDS.
constexpr uint32_t FIELD_TILE64 = 256;

__global__ void __launch_bounds__(FIELD_BLOCK) k_update_e_b_f64(FieldArgs a) {
    __shared__ double4 s_pq[FIELD_TILE64];
    __shared__ double4 s_vd[FIELD_TILE64];
    const uint32_t tid = threadIdx.x;
    const uint32_t NX = a.nx, NY = a.ny, NZ = a.nz;
    const uint32_t nd = 1u << a.D;
    const uint32_t lbx = NX / nd, lby = NY / nd, lbz = NZ / nd;
    const uint64_t n64 = (uint64_t)blockIdx.x * FIELD_BLOCK + tid;
    bool act = n64 < a.N;
    const uint32_t n = act ? (uint32_t)n64 : 0u;
    if (act) act = (a.flags[n] & 0x1F) != 0x01;
    const uint32_t z = n / (NX * NY), t = n - z * NX * NY, y = t / NX, x = t - y * NX;
    const double px = (double)(float)x, py = (double)(float)y, pz = (double)(float)z;
    const uint32_t own = x / lbx + (y / lby + z / lbz * nd) * nd;
    double ex = 0.0, ey = 0.0, ez = 0.0, bx = 0.0, by = 0.0, bz = 0.0;
    const double* q = (const double*)a.q;
    const double* u = (const double*)a.u;
    if (a.near && act) {
        const uint32_t x0 = (x / a.dsx) * a.dsx, y0 = (y / a.dsy) * a.dsy, z0 = (z / a.dsz) * a.dsz;
        const uint32_t xu = min(x0 + a.dsx, NX), yu = min(y0 + a.dsy, NY), zu = min(z0 + a.dsz, NZ);
        for (uint32_t xx = x0; xx < xu; xx++)
            for (uint32_t yy = y0; yy < yu; yy++)
                for (uint32_t zz = z0; zz < zu; zz++) {
                    const uint32_t nc = xx + (yy + zz * NY) * NX;
                    if (nc == n) continue;
                    const double qc = q[nc];
                    if (qc == 0.0) continue;
                    interact_exact<double>(qc, u[nc], u[(uint64_t)nc + a.N], u[(uint64_t)nc + 2 * a.N],
                                           px - (double)(float)xx, py - (double)(float)yy, pz - (double)(float)zz,
                                           ex, ey, ez, bx, by, bz);
                }
    }
    const double4* spq = (const double4*)a.src_pq;
    const double4* svd = (const double4*)a.src_vd;
    const uint32_t count = *a.src_count;
    for (uint32_t base = 0; base < count; base += FIELD_TILE64) {
        const uint32_t m = min(FIELD_TILE64, count - base);
        __syncthreads();
        for (uint32_t i = tid; i < m; i += FIELD_BLOCK) {
            s_pq[i] = spq[base + i];
            s_vd[i] = svd[base + i];
        }
        __syncthreads();
        for (uint32_t s = 0; s < m; s++) {
            const double4 pq = s_pq[s];
            const double4 vd = s_vd[s];
            if ((uint32_t)vd.w == own) continue;
            interact_exact<double>(pq.w, vd.x, vd.y, vd.z, px - pq.x, py - pq.y, pz - pq.z, ex, ey, ez, bx, by, bz);
        }
    }
    if (!act) return;
    const uint64_t N = a.N;
    double* ed = (double*)a.e_dyn;
    double* bd = (double*)a.b_dyn;
    const double* es = (const double*)a.e_stat;
    const double* bs = (const double*)a.b_stat;
    const double ke = (double)a.ke, kmu = (double)a.kmu;
    ed[n] = es[n] + ke * ex;
    ed[N + n] = es[N + n] + ke * ey;
    ed[2 * N + n] = es[2 * N + n] + ke * ez;
    bd[n] = bs[n] + kmu * bx;
    bd[N + n] = bs[N + n] + kmu * by;
    bd[2 * N + n] = bs[2 * N + n] + kmu * bz;
}

// The field update on LOD-aligned tiles, FAST or EXACT arithmetic. A block covers
// (ROW*TX) x TY x TZ cells inside one top-level LOD block, so the own-LOD
// skip is the same for every cell of the block: a wave-uniform branch on the
// source's LOD index. Each thread owns ROW consecutive cells of one row: the
// source terms that depend only on y and z (r_y, r_z, r_y^2 + r_z^2 and the
// x component of v x r) are computed once per source and shared by the ROW
// cells. The sources are wave-uniform: they are read straight from the
// compacted list at a uniform index, which the compiler turns into scalar
// loads (SGPR operands for the VALU, no LDS staging and no barriers in the
// loop). A wavefront whose cells are all solid skips the sweep (its cells
// are not written). EXACT keeps the reference's arithmetic per interaction
// (only r_y^2 and r_z^2 are shared, the sum order is unchanged) with the
// specialised correctly rounded operators of cr_math.hpp: |r|^2 is a
// multiple of 1/4 in [1/4, 2^24) (cell minus block centre), |r|^3 lies in
// [1/8, 2^36] and the dividends r_i are 0 or in [1/2, 2^11].
// Held to 80 VGPRs (6 wavefronts per SIMD, no spills; the sweep's VALU
// instructions are the same as at 105 VGPRs) so that the scalar loads at the
// top of each source's iteration hide behind the other wavefronts.
template <int ROW, bool FAST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_update_e_b_tiled(FieldArgs a, uint32_t TX, uint32_t TY, uint32_t TZ) {
    __shared__ float s_epi[ROW * 256];
    const uint32_t NX = a.nx, NY = a.ny;
    const uint32_t tx = threadIdx.x % TX, ty = (threadIdx.x / TX) % TY, tz = threadIdx.x / (TX * TY);
    const uint32_t tiles_x = NX / (ROW * TX), tiles_y = NY / TY;
    const uint32_t b = blockIdx.x;
    const uint32_t bx = b % tiles_x, by = (b / tiles_x) % tiles_y, bz = b / (tiles_x * tiles_y);
    const uint32_t x0 = bx * ROW * TX + ROW * tx, y = by * TY + ty, z = bz * TZ + tz;
    const uint32_t nd = 1u << a.D;
    const uint32_t own = (bx * ROW * TX) / (NX / nd) + ((by * TY) / (NY / nd) + (bz * TZ) / (a.nz / nd) * nd) * nd;
    const float py = (float)y, pz = (float)z;
    float px[ROW], ex[ROW], ey[ROW], ez[ROW], bxa[ROW], bya[ROW], bza[ROW];
#pragma unroll
    for (int k = 0; k < ROW; k++) {
        px[k] = (float)(x0 + k);
        ex[k] = ey[k] = ez[k] = bxa[k] = bya[k] = bza[k] = 0.0f;
    }
    // does any of this work-item's cells need the field?
    const uint8_t* const fl = a.flags + (uint64_t)x0 + ((uint64_t)y + (uint64_t)z * NY) * NX;
    bool fluid = false;
#pragma unroll
    for (int k = 0; k < ROW; k++) fluid |= (fl[k] & 0x1F) != 0x01;
    const float4* __restrict__ const spq = a.src_pq;
    const float4* __restrict__ const svd = a.src_vd;
    const uint32_t count = *a.src_count;
    // The sources are in ascending LOD index: the own LOD's position (if it
    // has a charge) is found by a bisection, and the sweep runs in two
    // branch-free ranges around it (the reference skips it, sim_kernels.cl:950).
    uint32_t lo = 0, hi = count;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (__float_as_uint(svd[mid].w) < own) lo = mid + 1;
        else hi = mid;
    }
    const uint32_t skip = (lo < count && __float_as_uint(svd[lo].w) == own) ? lo : count;
    auto sweep = [&](uint32_t s0, uint32_t s1) {
        if (s0 >= s1) return;
        // the next source's scalar loads are issued before this one's
        // arithmetic, so their latency hides behind it
        float4 vd_next = svd[s0], pq_next = spq[s0];
        // one source per iteration: the compiler then lays the ROW cells'
        // chains side by side (measured at 256^3, FAST: 17.0 ms against
        // 18.1 ms with two sources per iteration and 17.5 ms with four)
#pragma unroll 1
        for (uint32_t s = s0; s < s1; s++) {
            const float4 vd = vd_next, pq = pq_next;
            const uint32_t sn = s + 1 < s1 ? s + 1 : s;
            vd_next = svd[sn];
            pq_next = spq[sn];
            const float ry = py - pq.y, rz = pz - pq.z;
            if constexpr (FAST) {
                const float syz = __builtin_fmaf(ry, ry, rz * rz);
                const float cx = __builtin_fmaf(vd.y, rz, -(vd.z * ry));  // (v x r).x
                const float m1 = vd.x * rz, m2 = vd.x * ry;
#pragma unroll
                for (int k = 0; k < ROW; k++) {
                    const float rx = px[k] - pq.x;
                    const float inv = __builtin_amdgcn_rsqf(__builtin_fmaf(rx, rx, syz));
                    const float t = pq.w * (inv * inv * inv);
                    ex[k] = __builtin_fmaf(t, rx, ex[k]);
                    ey[k] = __builtin_fmaf(t, ry, ey[k]);
                    ez[k] = __builtin_fmaf(t, rz, ez[k]);
                    bxa[k] = __builtin_fmaf(t, cx, bxa[k]);
                    bya[k] = __builtin_fmaf(t, __builtin_fmaf(vd.z, rx, -m1), bya[k]);  // (v x r).y
                    bza[k] = __builtin_fmaf(t, __builtin_fmaf(-vd.y, rx, m2), bza[k]);  // (v x r).z
                }
            } else {
                // interact_exact with r_y^2, r_z^2 hoisted: pre = r / |r|^3,
                // e += q pre, b += q (v x pre), every operation rounded on its own
                const float ry2 = ry * ry, rz2 = rz * rz;
                const float q = pq.w;
#pragma unroll
                for (int k = 0; k < ROW; k++) {
                    const float rx = px[k] - pq.x;
                    const float sm = sqrt_cr((rx * rx + ry2) + rz2);
                    const float cbm = (sm * sm) * sm;
                    const float prx = div_cr(rx, cbm), pry = div_cr(ry, cbm), prz = div_cr(rz, cbm);
                    ex[k] = ex[k] + q * prx;
                    ey[k] = ey[k] + q * pry;
                    ez[k] = ez[k] + q * prz;
                    const float crx = vd.y * prz - vd.z * pry, cry = vd.z * prx - vd.x * prz,
                                crz = vd.x * pry - vd.y * prx;
                    bxa[k] = bxa[k] + q * crx;
                    bya[k] = bya[k] + q * cry;
                    bza[k] = bza[k] + q * crz;
                }
            }
        }
    };
    if (__any(fluid)) {
        sweep(0, skip);
        sweep(skip + 1, count);
    }
    // Epilogue through LDS, one component at a time: each work-item's ROW
    // consecutive cells go to LDS in block order, then the block writes whole
    // row segments (ROW*TX cells) with consecutive lanes on consecutive cells,
    // so that every store (and E_stat / B_stat load) is contiguous. Solid
    // cells (among them every cell of a wavefront that skipped the sweep) are
    // not written.
    const uint64_t N = a.N;
    const uint32_t W = ROW * TX;
    const uint32_t own0 = x0 - bx * W + W * (ty + TY * tz);  // first own cell in block order
    uint32_t live = 0;  // bit j: cell threadIdx.x + j * blockDim.x is not solid
    uint64_t gn[ROW];
#pragma unroll
    for (int j = 0; j < ROW; j++) {
        const uint32_t i = threadIdx.x + j * blockDim.x;
        const uint32_t lx = i % W, ly = (i / W) % TY, lz = i / (W * TY);
        gn[j] = (uint64_t)(bx * W + lx) + ((uint64_t)(by * TY + ly) + (uint64_t)(bz * TZ + lz) * NY) * NX;
        if ((a.flags[gn[j]] & 0x1F) != 0x01) live |= 1u << j;
    }
    float jd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int c = 0; c < 6; c++) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ROW; k++) {
            const float v = c == 0 ? ex[k] : c == 1 ? ey[k] : c == 2 ? ez[k] : c == 3 ? bxa[k] : c == 4 ? bya[k] : bza[k];
            s_epi[own0 + k] = v;
        }
        __syncthreads();
        const uint64_t off = (uint64_t)(c % 3) * N;
#pragma unroll
        for (int j = 0; j < ROW; j++) {
            if (!(live >> j & 1u)) continue;
            const float v = s_epi[threadIdx.x + j * blockDim.x];
            const uint64_t n = gn[j];
            if (a.raw != nullptr) {
                a.raw[(uint64_t)c * N + n] = v;
                continue;
            }
            float* const dyn = c < 3 ? a.e_dyn : a.b_dyn;
            if (a.judge != nullptr) {  // the surrogate's raw sum, about to be replaced
                const float sur = dyn[off + n];
                jd[c < 3 ? 0 : 2] = __builtin_fmaf(sur - v, sur - v, jd[c < 3 ? 0 : 2]);
                jd[c < 3 ? 1 : 3] = __builtin_fmaf(v, v, jd[c < 3 ? 1 : 3]);
            }
            if (c < 3) dyn[off + n] = a.e_stat[off + n] + a.ke * v;
            else dyn[off + n] = a.b_stat[off + n] + a.kmu * v;
        }
    }
    if (a.judge != nullptr) judge_block_sums(jd, a.judge);
}

__global__ void k_combine_fields(FieldArgs a) {
    const uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t N = a.N;
    if (n >= N || (a.flags[n] & 0x1F) == 0x01) return;
    const float* r = a.raw;
    a.e_dyn[n] = a.e_stat[n] + a.ke * r[n];
    a.e_dyn[N + n] = a.e_stat[N + n] + a.ke * r[N + n];
    a.e_dyn[2 * N + n] = a.e_stat[2 * N + n] + a.ke * r[2 * N + n];
    a.b_dyn[n] = a.b_stat[n] + a.kmu * r[3 * N + n];
    a.b_dyn[N + n] = a.b_stat[N + n] + a.kmu * r[4 * N + n];
    a.b_dyn[2 * N + n] = a.b_stat[2 * N + n] + a.kmu * r[5 * N + n];
}

// Largest power of two <= cap dividing v.
static uint32_t pow2_div(uint32_t v, uint32_t cap) {
    uint32_t p = 1;
    while (p * 2 <= cap && v % (p * 2) == 0) p *= 2;
    return p;
}

// Work-items per tile along x, y, z for LOD blocks of wx work-items (cells /
// ROW) by ly by lz cells: powers of two where they give the most work-items,
// else divisors with a product that is a multiple of 64 (whole wavefronts),
// at most 16 along x and 256 in all; the larger x extent wins a tie.
static void tile_shape(uint32_t wx, uint32_t ly, uint32_t lz, uint32_t& TX, uint32_t& TY, uint32_t& TZ) {
    TX = pow2_div(wx, 16);
    TY = pow2_div(ly, 256 / TX);
    TZ = pow2_div(lz, 256 / (TX * TY));
    uint32_t best = TX * TY * TZ;
    for (uint32_t tx = 16; tx >= 1; tx--) {
        if (wx % tx) continue;
        for (uint32_t ty = 256 / tx; ty >= 1; ty--) {
            if (ly % ty) continue;
            for (uint32_t tz = 256 / (tx * ty); tz >= 1; tz--) {
                const uint32_t p = tx * ty * tz;
                if (lz % tz || p % 64u || p <= best) continue;
                TX = tx; TY = ty; TZ = tz; best = p;
            }
        }
    }
}

// Float64 sum of x over fluid cells (flags & 0x1F != 0x01): one partial per
// block, summed on the host in block order.
template <class T>
__global__ void __launch_bounds__(256) k_sum_f64(const T* __restrict__ x, const uint8_t* __restrict__ flags,
                                                 uint64_t n, double* __restrict__ partial) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if ((flags[i] & 0x1F) != 0x01) acc += (double)x[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t s = 128; s >= 1; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

hipError_t launch_prep_sources(const void* qu_lod, bool f64, uint32_t lo, uint32_t hi, uint32_t nx, uint32_t ny,
                               uint32_t nz, uint32_t D, void* src_pq, void* src_vd, uint32_t* count, hipStream_t s) {
    const uint32_t nd = 1u << D;
    const float dx = (float)(nx / nd), dy = (float)(ny / nd), dz = (float)(nz / nd);
    if (f64)
        hipLaunchKernelGGL(k_prep_sources<double>, dim3(1), dim3(1024), 0, s, (const double4*)qu_lod, lo, hi, nd,
                           (double)dx, (double)dy, (double)dz, (double4*)src_pq, (double4*)src_vd, count);
    else
        hipLaunchKernelGGL(k_prep_sources<float>, dim3(1), dim3(1024), 0, s, (const float4*)qu_lod, lo, hi, nd, dx,
                           dy, dz, (float4*)src_pq, (float4*)src_vd, count);
    return hipGetLastError();
}

// Launch shape of the field update: the LOD-tiled kernel with ROW cells per
// work-item and TX x TY x TZ work-items per block (row 0: the generic kernel).
struct FieldShape {
    uint32_t row, TX, TY, TZ, blocks;
};
static FieldShape field_shape(const FieldArgs& a, bool fast) {
    constexpr int CPT_EXACT = 1, CPT_FAST = 2;
    if (!a.f64 && !a.near && a.D > 0) {
        // LOD-aligned tiles: (ROW*TX) | lod block x, TY | lod block y, TZ | lod block z
        const uint32_t nd = 1u << a.D;
        const uint32_t lbx = a.nx / nd, lby = a.ny / nd, lbz = a.nz / nd;
        for (uint32_t row : {8u, 4u}) {
            if (lbx % row) continue;
            uint32_t TX, TY, TZ;
            tile_shape(lbx / row, lby, lbz, TX, TY, TZ);
            if (TX * TY * TZ < 64) continue;
            return FieldShape{row, TX, TY, TZ, (a.nx / (row * TX)) * (a.ny / TY) * (a.nz / TZ)};
        }
    }
    const uint64_t per = (uint64_t)FIELD_BLOCK * (a.f64 ? 1 : (fast ? CPT_FAST : CPT_EXACT));
    return FieldShape{0, 0, 0, 0, (uint32_t)((a.N + per - 1) / per)};
}

uint32_t field_blocks(const FieldArgs& a, bool fast) { return field_shape(a, fast).blocks; }

hipError_t launch_update_e_b(const FieldArgs& a, bool fast, hipStream_t s) {
    constexpr int CPT_EXACT = 1, CPT_FAST = 2;
    const FieldShape sh = field_shape(a, fast);
    if (a.f64) {
        if (a.judge != nullptr) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_update_e_b_f64, dim3(sh.blocks), dim3(FIELD_BLOCK), 0, s, a);
        return hipGetLastError();
    }
    if (sh.row) {
        const uint32_t TX = sh.TX, TY = sh.TY, TZ = sh.TZ;
        const dim3 g(sh.blocks), bl(TX * TY * TZ);
        if (fast) {
            if (sh.row == 8) hipLaunchKernelGGL((k_update_e_b_tiled<8, true>), g, bl, 0, s, a, TX, TY, TZ);
            else hipLaunchKernelGGL((k_update_e_b_tiled<4, true>), g, bl, 0, s, a, TX, TY, TZ);
        } else {
            if (sh.row == 8) hipLaunchKernelGGL((k_update_e_b_tiled<8, false>), g, bl, 0, s, a, TX, TY, TZ);
            else hipLaunchKernelGGL((k_update_e_b_tiled<4, false>), g, bl, 0, s, a, TX, TY, TZ);
        }
        return hipGetLastError();
    }
    const dim3 g(sh.blocks), b(FIELD_BLOCK);
    if (fast) {
        if (a.near) hipLaunchKernelGGL((k_update_e_b<true, true, CPT_FAST>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_update_e_b<true, false, CPT_FAST>), g, b, 0, s, a);
    } else {
        if (a.near) hipLaunchKernelGGL((k_update_e_b<false, true, CPT_EXACT>), g, b, 0, s, a);
        else hipLaunchKernelGGL((k_update_e_b<false, false, CPT_EXACT>), g, b, 0, s, a);
    }
    return hipGetLastError();
}

// Judge partials [nblocks][4] -> out[4]: thread t sums the blocks t, t + 256,
// ... of component t % 4 group, then the 64 group sums of each component are
// added in order (a fixed order: the same residual on every run).
__global__ void __launch_bounds__(256) k_judge_reduce(const double* __restrict__ partial, uint32_t nblocks,
                                                      double* __restrict__ out) {
    __shared__ double sh[256];
    const uint32_t t = threadIdx.x, comp = t & 3u, grp = t >> 2;
    double acc = 0.0;
    for (uint32_t b = grp; b < nblocks; b += 64u) acc += partial[(size_t)b * 4 + comp];
    sh[t] = acc;
    __syncthreads();
    if (t < 4) {
        double s = 0.0;
        for (uint32_t g = 0; g < 64u; g++) s += sh[g * 4 + t];
        out[t] = s;
    }
}

hipError_t launch_judge_reduce(const double* partial, uint32_t nblocks, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_judge_reduce, dim3(1), dim3(256), 0, s, partial, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_combine_fields(const FieldArgs& a, hipStream_t s) {
    if (a.raw == nullptr) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_combine_fields, dim3((uint32_t)((a.N + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sum_f64(const void* x, bool f64, const uint8_t* flags, uint64_t n, double* partial,
                          uint32_t nblocks, hipStream_t s) {
    if (f64) hipLaunchKernelGGL(k_sum_f64<double>, dim3(nblocks), dim3(256), 0, s, (const double*)x, flags, n, partial);
    else hipLaunchKernelGGL(k_sum_f64<float>, dim3(nblocks), dim3(256), 0, s, (const float*)x, flags, n, partial);
    return hipGetLastError();
}

}  // namespace ion
