// Mesh voxelisation and the static-field precompute for gfx950.
//
// voxelize_mesh (src/kernels/sim_kernels.cl:1150-1231): one thread per ray
// column, triangles staged through LDS in tiles shared by the workgroup,
// the per-column intersection distances kept in LDS (not in per-lane
// scratch). psi_from_mesh / static_e_from_mesh (:1234-1251, :1278-1299)
// loop over every lattice cell and test its flag; here the source cells are
// first compacted, in ascending cell order, into a list (count, scan,
// scatter), so the O(N^2) flag scan becomes N x sources interactions over
// LDS-staged source tiles, summed per target in the same order as the
// reference. All arithmetic is the reference's, unfused and correctly
// rounded (the library is built with -ffp-contract=off and correctly rounded
// division and square root).
#include "cr_math.hpp"
#include "mesh.hpp"

namespace ion {
namespace mesh {
namespace {

constexpr uint32_t VOX_BLOCK = 64;   // one wavefront per workgroup
constexpr uint32_t VOX_TILE = 128;   // triangles per LDS tile

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

__global__ __launch_bounds__(VOX_BLOCK) void k_voxelize(VoxArgs a) {
    __shared__ float tri_s[VOX_TILE * 9];
    __shared__ uint16_t dist_s[MAX_HITS][VOX_BLOCK];
    const uint32_t tid = threadIdx.x;
    const uint32_t col = blockIdx.x * VOX_BLOCK + tid;
    const uint32_t NX = a.nx, NY = a.ny, NZ = a.nz;
    const uint32_t dir = a.direction;
    // column start cell (sim_kernels.cl:1163-1168); DEF_OX = DEF_OY = DEF_OZ = 0 on one domain
    uint32_t x, y, z;
    if (dir == 0u) {
        x = (uint32_t)clampi((int)a.x0, 0, (int)NX - 1); y = col % NY; z = col / NY;
    } else if (dir == 1u) {
        x = col / NZ; y = (uint32_t)clampi((int)a.y0, 0, (int)NY - 1); z = col % NZ;
    } else {
        x = col % NX; y = col / NX; z = (uint32_t)clampi((int)a.z0, 0, (int)NZ - 1);
    }
    // ray origin: position(xyz) + offset (:131-133, :1169-1170)
    const float ox = 0.5f * (float)(int)NX - 0.5f, oy = 0.5f * (float)(int)NY - 0.5f, oz = 0.5f * (float)(int)NZ - 0.5f;
    const float rx = ((float)x + 0.5f - 0.5f * (float)NX) + ox;
    const float ry = ((float)y + 0.5f - 0.5f * (float)NY) + oy;
    const float rz = ((float)z + 0.5f - 0.5f * (float)NZ) + oz;
    const float rdx = (float)(dir == 0u), rdy = (float)(dir == 1u), rdz = (float)(dir == 2u);
    bool outside;
    if (dir == 0u) outside = ry < a.y0 || rz < a.z0 || ry >= a.y1 || rz >= a.z1;
    else if (dir == 1u) outside = rx < a.x0 || rz < a.z0 || rx >= a.x1 || rz >= a.z1;
    else outside = rx < a.x0 || ry < a.y0 || rx >= a.x1 || ry >= a.y1;
    const bool active = col < a.A && !outside;
    if (!__syncthreads_or(active)) return;  // workgroup-uniform

    uint32_t hits = 0u, back = 0u;
    for (uint32_t t0 = 0; t0 < a.T; t0 += VOX_TILE) {
        const uint32_t nt = min(VOX_TILE, a.T - t0);
        __syncthreads();
        for (uint32_t k = tid; k < nt * 9u; k += VOX_BLOCK) tri_s[k] = a.tri[(uint64_t)t0 * 9u + k];
        __syncthreads();
        if (!active) continue;
        for (uint32_t i = 0; i < nt; i++) {
            const float* p = tri_s + 9u * i;
            // bidirectional Moeller-Trumbore ray-triangle test (:1182-1185)
            const float ux = p[3] - p[0], uy = p[4] - p[1], uz = p[5] - p[2];
            const float vx = p[6] - p[0], vy = p[7] - p[1], vz = p[8] - p[2];
            const float wx = rx - p[0], wy = ry - p[1], wz = rz - p[2];
            const float hx = rdy * vz - rdz * vy, hy = rdz * vx - rdx * vz, hz = rdx * vy - rdy * vx;
            const float qx = wy * uz - wz * uy, qy = wz * ux - wx * uz, qz = wx * uy - wy * ux;
            const float f = 1.0f / (ux * hx + uy * hy + uz * hz);
            const float s = f * (wx * hx + wy * hy + wz * hz);
            const float t = f * (rdx * qx + rdy * qy + rdz * qz);
            const float d = f * (vx * qx + vy * qy + vz * qz);
            if (s >= 0.0f && s < 1.0f && t >= 0.0f && s + t < 1.0f) {
                if (d > 0.0f) {  // ahead of the column start
                    if (hits < MAX_HITS && d < 65536.0f) dist_s[hits][tid] = (uint16_t)(uint32_t)d;
                    hits++;
                } else {
                    back++;  // behind: second ray for the inside test
                }
            }
        }
    }
    if (!active) return;
    // a hit counted but not stored (d >= 65536) reads as 65535 (undefined in the reference)
    const uint32_t m = min(hits, MAX_HITS);
    // insertion sort of the stored distances (:1192-1201)
    for (uint32_t i = 1; i < m; i++) {
        const uint16_t v = dist_s[i][tid];
        int j = (int)i - 1;
        while (j >= 0 && dist_s[j][tid] > v) {
            dist_s[j + 1][tid] = dist_s[j][tid];
            j--;
        }
        dist_s[j + 1][tid] = v;
    }
    // walk the column, toggling inside/outside at each intersection (:1203-1229)
    bool inside = (hits & 1u) && (back & 1u);
    uint32_t it = ((hits & 1u) != (back & 1u)) ? 1u : 0u;
    uint32_t h0, hmax;
    if (dir == 0u) { h0 = x; hmax = (uint32_t)clampi((int)a.x1, 0, (int)NX); }
    else if (dir == 1u) { h0 = y; hmax = (uint32_t)clampi((int)a.y1, 0, (int)NY); }
    else { h0 = z; hmax = (uint32_t)clampi((int)a.z1, 0, (int)NZ); }
    const uint32_t hmesh = hits ? h0 + (uint32_t)dist_s[min(hits - 1u, MAX_HITS - 1u)][tid] : 0u;
    const uint64_t N = a.N;
    for (uint32_t h = h0; h < hmax; h++) {
        while (it < hits && h > h0 + (uint32_t)dist_s[min(it, MAX_HITS - 1u)][tid]) {
            inside = !inside;
            it++;
        }
        inside = inside && (it < hits && h < hmesh);
        if (inside) {
            const uint32_t cx = dir == 0u ? h : x, cy = dir == 1u ? h : y, cz = dir == 2u ? h : z;
            const uint64_t n = (uint64_t)cx + ((uint64_t)cy + (uint64_t)cz * NY) * NX;
            a.flags[n] = (uint8_t)((a.flags[n] & (uint8_t)~TYPE_BO) | (uint8_t)a.flag);
            if (a.b_dyn) {
                if (a.flag & TYPE_M) {  // magnetised solid: magnetisation into B_dyn
                    a.b_dyn[n] = a.mx;
                    a.b_dyn[N + n] = a.my;
                    a.b_dyn[2 * N + n] = a.mz;
                } else if ((a.flag & TYPE_F) || (a.flag & TYPE_C)) {  // charged solid: charge into B_dyn.x
                    a.b_dyn[n] = a.mx;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Source compaction in ascending cell order: count per block, one-block
// exclusive scan of the counts, then an ordered scatter.
// ---------------------------------------------------------------------------
constexpr uint32_t SEL_BLOCK = 256;
constexpr uint32_t SEL_ROUNDS = 16;
constexpr uint32_t SEL_CELLS = SEL_BLOCK * SEL_ROUNDS;  // cells per block
constexpr uint32_t SCAN_THREADS = 1024;

// Magnetisation components for which k_psi may use the specialised division:
// 0 or 2^-40 <= |v| <= 2^40 (finite).
__device__ __forceinline__ bool in_div_range(float v) {
    const float a = fabsf(v);
    return a == 0.0f || (a >= 0x1p-40f && a <= 0x1p40f);
}

__device__ __forceinline__ bool selected(int mode, uint8_t f) {
    return mode == 0 ? (f & TYPE_M) != 0 : ((f & TYPE_F) != 0 || (f & TYPE_C) != 0);
}

__global__ __launch_bounds__(SEL_BLOCK) void k_select_count(int mode, const uint8_t* flags, uint64_t N,
                                                            uint32_t* block_counts) {
    __shared__ uint32_t part[SEL_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SEL_CELLS;
    uint32_t c = 0;
    for (uint32_t r = 0; r < SEL_ROUNDS; r++) {
        const uint64_t n = base + (uint64_t)r * SEL_BLOCK + threadIdx.x;
        if (n < N && selected(mode, flags[n])) c++;
    }
    // wavefront sum, then across the 4 wavefronts
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < SEL_BLOCK / 64; w++) t += part[w];
        block_counts[blockIdx.x] = t;
    }
}

// Exclusive scan of nb block counts in place; block_counts[nb] = total.
__global__ __launch_bounds__(SCAN_THREADS) void k_select_scan(uint32_t* block_counts, uint32_t nb, uint32_t* total) {
    __shared__ uint32_t sums[SCAN_THREADS];
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (nb + SCAN_THREADS - 1) / SCAN_THREADS;
    const uint32_t lo = min(nb, tid * chunk), hi = min(nb, lo + chunk);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; i++) s += block_counts[i];
    sums[tid] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan of the per-thread sums
    for (uint32_t o = 1; o < SCAN_THREADS; o <<= 1) {
        const uint32_t v = tid >= o ? sums[tid - o] : 0u;
        __syncthreads();
        sums[tid] += v;
        __syncthreads();
    }
    uint32_t run = tid ? sums[tid - 1] : 0u;
    for (uint32_t i = lo; i < hi; i++) {
        const uint32_t c = block_counts[i];
        block_counts[i] = run;
        run += c;
    }
    if (tid == SCAN_THREADS - 1) {
        block_counts[nb] = sums[tid];
        *total = sums[tid];
    }
}

__global__ __launch_bounds__(SEL_BLOCK) void k_select_scatter(int mode, const uint8_t* flags, const float* m,
                                                              uint64_t N, uint32_t nx, uint32_t ny,
                                                              const uint32_t* block_offsets, Sources out) {
    __shared__ uint32_t wsum[SEL_BLOCK / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * SEL_CELLS;
    uint32_t run = block_offsets[blockIdx.x];
    const uint32_t NXY = nx * ny;
    for (uint32_t r = 0; r < SEL_ROUNDS; r++) {
        const uint64_t n = base + (uint64_t)r * SEL_BLOCK + tid;
        const bool p = n < N && selected(mode, flags[n]);
        const uint64_t mask = __ballot(p);
        const uint32_t rank = __popcll(mask & ((1ull << lane) - 1ull));
        __syncthreads();
        if (lane == 0) wsum[wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t before = run;
        for (uint32_t w = 0; w < wave; w++) before += wsum[w];
        if (p) {
            const uint32_t k = before + rank;
            const uint32_t nn = (uint32_t)n;  // coordinates(n), sim_kernels.cl:134-137
            const uint32_t t = nn % NXY;
            const uint32_t cx = t % nx, cy = t / nx, cz = nn / NXY;
            const uint32_t off = mode == 0 ? 1u : 0u;
            out.pos[k] = make_float4((float)(cx + off), (float)(cy + off), (float)(cz + off), 0.0f);
            if (mode == 0) {
                const float mx = m[n], my = m[N + n], mz = m[2 * N + n];
                out.val[k] = make_float4(mx, my, mz, 0.0f);
                if (!(in_div_range(mx) && in_div_range(my) && in_div_range(mz))) atomicOr(out.count + 1, 1u);
            } else {
                out.val[k] = make_float4(m[n], 0.0f, 0.0f, 0.0f);
            }
        }
        for (uint32_t w = 0; w < SEL_BLOCK / 64; w++) run += wsum[w];
    }
}

// ---------------------------------------------------------------------------
// psi_from_mesh (sim_kernels.cl:1234-1251) over compacted magnet sources.
//
// The operands keep the specialised correctly rounded operators of
// cr_math.hpp exact: |r|^2 is an exact integer in [1, 2^24) (cell offsets),
// |r|^3 lies in [1, 2^36], and with every magnetisation component 0 or
// within [2^-40, 2^40] the dividend r.M is 0 or within [2^-63, 2^54], the
// quotient normal. The scatter pass checks the magnetisation range and k_psi
// falls back to the general operators otherwise.
// ---------------------------------------------------------------------------
constexpr uint32_t FLD_BLOCK = 256;
constexpr uint32_t FLD_TILE = 256;

template <bool SPECIAL>
__device__ __forceinline__ float psi_tile(const float4* sp, const float4* sv, uint32_t m, float cx, float cy, float cz,
                                          float psic) {
    for (uint32_t j = 0; j < m; j++) {
        const float4 p = sp[j], v = sv[j];
        const float dx = cx - p.x, dy = cy - p.y, dz = cz - p.z;
        const float l2 = dx * dx + dy * dy + dz * dz;
        const float d = dx * v.x + dy * v.y + dz * v.z;
        if (SPECIAL) {
            // the own cell (l = 0) adds +0, which leaves psic unchanged: psic
            // starts at +0 and a round-to-nearest sum is never -0 unless both
            // terms are
            const float l = sqrt_cr(l2);
            const float q = div_cr(d, l * l * l);
            psic += l2 != 0.0f ? q : 0.0f;
        } else {
            const float l = sqrtf(l2);
            if (!(l == 0.0f)) psic += d / (l * l * l);
        }
    }
    return psic;
}

__global__ __launch_bounds__(FLD_BLOCK) void k_psi(Sources src, float* psi, uint32_t nx, uint32_t ny, uint32_t nz) {
    __shared__ float4 sp[FLD_TILE];
    __shared__ float4 sv[FLD_TILE];
    const uint32_t NXp = nx + 2u, NYp = ny + 2u;
    const uint32_t Np = NXp * NYp * (nz + 2u);
    const uint32_t n = blockIdx.x * FLD_BLOCK + threadIdx.x;
    const uint32_t t = n % (NXp * NYp);  // coordinates_sl(n, nx + 2, ny + 2)
    const float cx = (float)(t % NXp), cy = (float)(t / NXp), cz = (float)(n / (NXp * NYp));
    const uint32_t cnt = src.count[0];
    // specialised operators need |r|^2 < 2^24: padded extents below 2^11 per axis
    const bool special = src.count[1] == 0u && NXp < 2048u && NYp < 2048u && nz + 2u < 2048u;
    float psic = 0.0f;
    for (uint32_t b0 = 0; b0 < cnt; b0 += FLD_TILE) {
        const uint32_t m = min(FLD_TILE, cnt - b0);
        __syncthreads();
        if (threadIdx.x < m) {
            sp[threadIdx.x] = src.pos[b0 + threadIdx.x];
            sv[threadIdx.x] = src.val[b0 + threadIdx.x];
        }
        __syncthreads();
        psic = special ? psi_tile<true>(sp, sv, m, cx, cy, cz, psic) : psi_tile<false>(sp, sv, m, cx, cy, cz, psic);
    }
    // psic / (4.0f * M_PI): M_PI is a double constant, so the quotient is taken in double
    if (n < Np) psi[n] = (float)((double)psic / (4.0 * 3.14159265358979323846));
}

// static_b_from_mesh (sim_kernels.cl:1253-1276).
__global__ __launch_bounds__(FLD_BLOCK) void k_static_b(const uint8_t* flags, float* B, const float* psi, uint32_t nx,
                                                        uint32_t ny, uint32_t nz, float kmu0) {
    const uint64_t N = (uint64_t)nx * ny * nz;
    const uint64_t n = (uint64_t)blockIdx.x * FLD_BLOCK + threadIdx.x;
    if (n >= N) return;
    if ((flags[n] & TYPE_S) == TYPE_S) return;
    const uint32_t NXY = nx * ny;
    const uint32_t t = (uint32_t)n % NXY;
    const uint32_t x = t % nx + 1u, y = t / nx + 1u, z = (uint32_t)n / NXY + 1u;
    const uint32_t l0 = nx + 2u, l1 = ny + 2u;
    const uint32_t c = x + (y + z * l1) * l0;
    const uint32_t yo = l0, zo = l0 * l1;
    const float gx = (psi[c + 1u] - psi[c - 1u]) / 2.0f;
    const float gy = (psi[c + yo] - psi[c - yo]) / 2.0f;
    const float gz = (psi[c + zo] - psi[c - zo]) / 2.0f;
    const float k = -kmu0;
    B[n] += k * gx;
    B[N + n] += k * gy;
    B[2 * N + n] += k * gz;
}

// static_e_from_mesh (sim_kernels.cl:1278-1299) over compacted charge sources.
__global__ __launch_bounds__(FLD_BLOCK) void k_static_e(Sources src, const uint8_t* flags, float* E, uint32_t nx,
                                                        uint32_t ny, uint32_t nz, float ke) {
    __shared__ float4 sp[FLD_TILE];
    const uint64_t N = (uint64_t)nx * ny * nz;
    const uint64_t n = (uint64_t)blockIdx.x * FLD_BLOCK + threadIdx.x;
    const bool live = n < N && (flags[n] & TYPE_S) != TYPE_S;
    const uint32_t NXY = nx * ny;
    const uint32_t t = (uint32_t)n % NXY;
    const float cx = (float)(t % nx), cy = (float)(t / nx), cz = (float)((uint32_t)n / NXY);
    const uint32_t cnt = *src.count;
    float ex = 0.0f, ey = 0.0f, ez = 0.0f;
    for (uint32_t b0 = 0; b0 < cnt; b0 += FLD_TILE) {
        const uint32_t m = min(FLD_TILE, cnt - b0);
        __syncthreads();
        if (threadIdx.x < m) {
            float4 p = src.pos[b0 + threadIdx.x];
            p.w = src.val[b0 + threadIdx.x].x;
            sp[threadIdx.x] = p;
        }
        __syncthreads();
        for (uint32_t j = 0; j < m; j++) {
            const float4 p = sp[j];
            const float dx = cx - p.x, dy = cy - p.y, dz = cz - p.z;
            const float l = sqrtf(dx * dx + dy * dy + dz * dz);
            if (!(l == 0.0f)) {
                const float c3 = l * l * l;
                ex += dx * p.w / c3;
                ey += dy * p.w / c3;
                ez += dz * p.w / c3;
            }
        }
    }
    if (!live) return;
    E[n] += ex * ke;
    E[N + n] += ey * ke;
    E[2 * N + n] += ez * ke;
}

}  // namespace

hipError_t launch_voxelize(const VoxArgs& a, hipStream_t s) {
    if (a.A == 0) return hipSuccess;
    hipLaunchKernelGGL(k_voxelize, dim3((a.A + VOX_BLOCK - 1) / VOX_BLOCK), dim3(VOX_BLOCK), 0, s, a);
    return hipGetLastError();
}

uint32_t select_blocks(uint64_t N) { return (uint32_t)((N + SEL_CELLS - 1) / SEL_CELLS); }

hipError_t launch_select_count(int mode, const uint8_t* flags, uint64_t N, uint32_t* block_counts, uint32_t* total,
                               hipStream_t s) {
    const uint32_t nb = select_blocks(N);
    hipLaunchKernelGGL(k_select_count, dim3(nb), dim3(SEL_BLOCK), 0, s, mode, flags, N, block_counts);
    hipLaunchKernelGGL(k_select_scan, dim3(1), dim3(SCAN_THREADS), 0, s, block_counts, nb, total);
    return hipGetLastError();
}

hipError_t launch_select_scatter(int mode, const uint8_t* flags, const float* m, uint64_t N, uint32_t nx, uint32_t ny,
                                 const uint32_t* block_counts, Sources out, hipStream_t s) {
    const uint32_t nb = select_blocks(N);
    hipLaunchKernelGGL(k_select_scatter, dim3(nb), dim3(SEL_BLOCK), 0, s, mode, flags, m, N, nx, ny, block_counts,
                       out);
    return hipGetLastError();
}

// every integer in [1, 2^24) and every quarter of one (|r|^2 of cell offsets
// from cell or LOD-block centres)
__global__ void k_check_sqrt(unsigned long long* mismatch, float scale) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x + 1u;  // 1 .. 2^24 - 1
    if (i >= (1u << 24)) return;
    const float x = (float)i * scale;
    if (__float_as_uint(sqrt_cr(x)) != __float_as_uint(sqrtf(x))) atomicAdd(mismatch, 1ull);
}

// pseudo-random (a, b) over the specialised ranges, with b = l^3 of an
// integer or quarter-integer l^2
__global__ void k_check_div(unsigned long long* mismatch, uint32_t seed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t h = i * 0x9E3779B9u ^ seed;
    h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
    uint32_t g = h * 0x2C1B3C6Du + 0x297A2D39u;
    g ^= g >> 15; g *= 0x27D4EB2Du; g ^= g >> 13;
    const float l2 = (float)(1u + (g % ((1u << 24) - 1u))) * ((seed & 2u) ? 0.25f : 1.0f);
    const float l = sqrtf(l2);
    const float b = l * l * l;
    // a: random sign, exponent in [-63, 54], random mantissa
    const uint32_t e = 127u - 63u + (h % 118u);
    const float a = __uint_as_float(((h << 8) & 0x80000000u) | (e << 23) | (g & 0x7FFFFFu));
    if (__float_as_uint(div_cr(a, b)) != __float_as_uint(a / b)) atomicAdd(mismatch, 1ull);
}

hipError_t run_psi_math_check(uint64_t* sqrt_mismatch, uint64_t* div_mismatch, uint64_t* div_samples) {
    unsigned long long* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, 2 * sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemset(d, 0, 2 * sizeof(unsigned long long));
    const uint32_t B = 256;
    const uint32_t rounds = 16, per = 1u << 24;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_check_sqrt, dim3(((1u << 24) + B - 1) / B), dim3(B), 0, 0, d, 1.0f);
        hipLaunchKernelGGL(k_check_sqrt, dim3(((1u << 24) + B - 1) / B), dim3(B), 0, 0, d, 0.25f);
        for (uint32_t r = 0; r < rounds; r++) hipLaunchKernelGGL(k_check_div, dim3(per / B), dim3(B), 0, 0, d + 1, r * 7919u + 1u);
        e = hipGetLastError();
    }
    unsigned long long h[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    *sqrt_mismatch = h[0];
    *div_mismatch = h[1];
    *div_samples = (uint64_t)rounds * per;
    return e;
}

hipError_t launch_psi(const Sources& src, float* psi, uint32_t nx, uint32_t ny, uint32_t nz, hipStream_t s) {
    const uint32_t Np = (nx + 2u) * (ny + 2u) * (nz + 2u);
    hipLaunchKernelGGL(k_psi, dim3((Np + FLD_BLOCK - 1) / FLD_BLOCK), dim3(FLD_BLOCK), 0, s, src, psi, nx, ny, nz);
    return hipGetLastError();
}

hipError_t launch_static_b(const uint8_t* flags, float* b, const float* psi, uint32_t nx, uint32_t ny, uint32_t nz,
                           float kmu0, hipStream_t s) {
    const uint64_t N = (uint64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_static_b, dim3((uint32_t)((N + FLD_BLOCK - 1) / FLD_BLOCK)), dim3(FLD_BLOCK), 0, s, flags,
                       b, psi, nx, ny, nz, kmu0);
    return hipGetLastError();
}

hipError_t launch_static_e(const Sources& src, const uint8_t* flags, float* e, uint32_t nx, uint32_t ny, uint32_t nz,
                           float ke, hipStream_t s) {
    const uint64_t N = (uint64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_static_e, dim3((uint32_t)((N + FLD_BLOCK - 1) / FLD_BLOCK)), dim3(FLD_BLOCK), 0, s, src,
                       flags, e, nx, ny, nz, ke);
    return hipGetLastError();
}

}  // namespace mesh
}  // namespace ion
