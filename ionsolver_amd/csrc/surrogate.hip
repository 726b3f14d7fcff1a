// Learned E/B field surrogate: a linear 4-level 3D convolutional neural
// operator on the level-of-detail (LOD) grid, with BF16 weights and
// activations and FP32 accumulation on the MFMA units.
//
// Input: the top-level LOD sums (charge q and current J = q v per LOD block,
// the data update_e_b_dynamic reads), output: E - E_stat and B - B_stat at
// every cell. All layers are convolutions, so every layer is one GEMM over an
// im2col matrix (bf16, K padded to 32): C[M][N] = A[M][K] . B[N][K]^T with
// v_mfma_f32_16x16x32_bf16. The output head is a sub-pixel convolution: for
// each LOD block it maps the 3^3 neighbourhood of level-0 features to the 6
// field components of every cell of the block, and its epilogue writes the
// fields (plus E_stat / B_stat) straight into E_dyn / B_dyn.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "surrogate.hpp"

#include <algorithm>

namespace ion {
namespace sur {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __bf16 f2bf(float x) { return (__bf16)x; }  // round to nearest even
__device__ __forceinline__ float bf2f(__bf16 x) { return (float)x; }

// ---------------------------------------------------------------------------
// GEMM: C[M][N] = alpha * A[M][K] . B[N][K]^T (+ C if accumulate), bf16
// operands with leading dimensions lda/ldb (elements), K % 32 == 0.
// 64 x 64 tile per 256-thread block, 4 waves of 32 x 32 (2 x 2 MFMA tiles).
// ---------------------------------------------------------------------------
// Lagrange basis of the three sample positions t_k = ((k + 1/2)/3) b - 1/2
// (spacing h = b/3) at offset u.
__device__ __forceinline__ void lagrange3(float u, float b, float L[3]) {
    const float h = b / 3.0f, t0 = 0.5f * h - 0.5f;
    const float s = (u - t0) / h;  // sample coordinate: nodes at 0, 1, 2
    L[0] = 0.5f * (s - 1.0f) * (s - 2.0f);
    L[1] = -s * (s - 2.0f);
    L[2] = 0.5f * s * (s - 1.0f);
}

// Far field of output column (component c, cell offset ox, oy, oz) of LOD
// block m: tri-quadratic interpolation of the block's 27 samples.
__device__ __forceinline__ float far_at(const GemmOut& o, uint32_t m, uint32_t c, uint32_t ox, uint32_t oy,
                                        uint32_t oz) {
    float Lx[3], Ly[3], Lz[3];
    lagrange3((float)ox, (float)o.bx, Lx);
    lagrange3((float)oy, (float)o.by, Ly);
    lagrange3((float)oz, (float)o.bz, Lz);
    const float* f = o.far + ((size_t)m * 6 + c) * FAR_PTS;
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) acc = __builtin_fmaf(Lz[i] * Ly[j] * Lx[k], f[i * 9 + j * 3 + k], acc);
    return acc;
}

constexpr int BM = 64, BN = 64, BK = 32, LDS_PAD = 8;

// Epilogue of one output element C[m][n] = v (alpha applied here).
template <int EPI>
__device__ __forceinline__ void epilogue(const GemmOut& out, int M, int N, int m, int n, float acc) {
    if (m >= M || n >= N) return;
    if (EPI == EPI_PARTIAL) {  // split K: this block's partial sum, reduced by k_splitk_reduce
        out.f32[((size_t)blockIdx.z * M + m) * N + n] = acc;
        return;
    }
    const float v = out.alpha * acc;
    if (EPI == EPI_F32) {
        float* p = out.f32 + (size_t)m * out.ldc + n;
        *p = out.accumulate ? *p + v : v;
    } else if (EPI == EPI_BF16) {
        out.bf16[(size_t)m * out.ldc + n] = f2bf(v);
    } else {  // EPI_HEAD: m = LOD block, n = component * cells_per_block + offset
        const uint32_t c = (uint32_t)n / out.cpb, o = (uint32_t)n - c * out.cpb;
        const uint32_t ox = o % out.bx, oy = (o / out.bx) % out.by, oz = o / (out.bx * out.by);
        const uint32_t X = (uint32_t)m % out.nd, Y = ((uint32_t)m / out.nd) % out.nd, Z = (uint32_t)m / (out.nd * out.nd);
        const uint64_t cell =
            (uint64_t)(X * out.bx + ox) + ((uint64_t)(Y * out.by + oy) + (uint64_t)(Z * out.bz + oz) * out.ny) * out.nx;
        if (out.fine != nullptr) {  // training: network-unit output (net + far field)
            const float far = far_at(out, (uint32_t)m, c, ox, oy, oz);
            out.fine[(size_t)c * out.ncell + cell] = acc + far * (out.gain / out.in_scale[c < 3 ? 0 : 1]);
        } else {
            if ((out.flags[cell] & 0x1F) == 0x01) return;
            const float far = far_at(out, (uint32_t)m, c, ox, oy, oz);
            const float s = c < 3 ? out.scale_e : out.scale_b;
            const float kf = c < 3 ? out.ke : out.kmu;
            const float* stat = c < 3 ? out.e_stat : out.b_stat;
            float* dyn = c < 3 ? out.e_dyn : out.b_dyn;
            const size_t idx = (size_t)(c % 3) * out.ncell + cell;
            const float in_s = out.in_scale[c < 3 ? 0 : 1];
            if (out.raw_out) dyn[idx] = in_s * acc / out.gain + far;
            else dyn[idx] = (stat[idx] + in_s * s * acc) + kf * far;
        }
    }
}

template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_nt(const __bf16* __restrict__ A, const __bf16* __restrict__ B, int M,
                                                 int N, int K, int lda, int ldb, GemmOut out) {
    __shared__ __bf16 As[BM][BK + LDS_PAD];
    __shared__ __bf16 Bs[BN][BK + LDS_PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    const int lr = tid >> 2, lk = (tid & 3) * 8;  // loader: row, k offset (8 bf16 = 16 bytes)
    // split K: block z takes the z-th of gridDim.z equal runs of K slices
    const int nk = K / BK, per = (nk + (int)gridDim.z - 1) / (int)gridDim.z;
    const int kb = (int)blockIdx.z * per * BK, ke = min(K, kb + per * BK);
    for (int k0 = kb; k0 < ke; k0 += BK) {
        bf16x8 va = (bf16x8){}, vb = (bf16x8){};
        if (m0 + lr < M) va = *(const bf16x8*)(A + (size_t)(m0 + lr) * lda + k0 + lk);
        if (n0 + lr < N) vb = *(const bf16x8*)(B + (size_t)(n0 + lr) * ldb + k0 + lk);
        __syncthreads();
        *(bf16x8*)&As[lr][lk] = va;
        *(bf16x8*)&Bs[lr][lk] = vb;
        __syncthreads();
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 fa[2], fb[2];
#pragma unroll
        for (int i = 0; i < 2; i++) fa[i] = *(const bf16x8*)&As[wm + i * 16 + fr][fk];
#pragma unroll
        for (int j = 0; j < 2; j++) fb[j] = *(const bf16x8*)&Bs[wn + j * 16 + fr][fk];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
    // lane holds C[(lane/16)*4 + r][lane%16] of each 16 x 16 tile
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                epilogue<EPI>(out, M, N, m0 + wm + i * 16 + (lane >> 4) * 4 + r, n0 + wn + j * 16 + (lane & 15),
                              acc[i][j][r]);
}

// 128 x 128 tile per 256-thread block, 4 waves of 64 x 64 (4 x 4 MFMA tiles);
// the next K slice is fetched into registers while the current one is
// multiplied out of one LDS stage, then stored into the other.
//
// Tile order (m_inner): when B has more tiles than A — the head, whose B
// (weights, N = 6 x cells per LOD block rows) is 2 GB at 1024^3 against
// 5.5 MB of A (one row per LOD block) — workgroups go to the 8 XCDs round
// robin (block b runs on XCD b % 8), and the blocks of one XCD walk its N
// tiles for a group of GM2 M tiles (1.4 MB of A, which stays in the XCD's
// 4 MB L2) before the next group: each B tile is fetched once per M group
// and served from L2 to the group's other M tiles, and A is fetched once
// per group instead of once per N tile. Otherwise N tiles are innermost.
constexpr int BM2 = 128, BN2 = 128;
constexpr uint32_t XCDS = 8, GM2 = 8;

// Held to 128 VGPRs (the accumulators in VGPRs, not AGPRs): 4 waves per SIMD.
template <int EPI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
    k_gemm_nt_128(const __bf16* __restrict__ A, const __bf16* __restrict__ B, int M, int N, int K, int lda, int ldb,
                  GemmOut out, int m_inner) {
    // The K-loop tiles (As, Bs) and the epilogue's staging (Cs) share one
    // LDS allocation: the epilogue's first barrier follows the last MFMA
    // reads. 32 KB per block, so LDS does not limit the blocks per CU
    // (registers allow 4). K-loop tiles: two stages of A and B, 128 x 32 bf16
    // each, unpadded; the 16-byte chunk c of row r sits at chunk
    // c ^ ((r >> 2) & 3), so that the 16-lane groups of both the tile stores
    // and the fragment reads touch 16 distinct bank groups. With two stages one
    // barrier per K slice suffices: stage k + 1 is stored while the others may
    // still read stage k.
    typedef __bf16 Tile[BM2][BK];
    typedef float CTile[4][16][64 + 4];
    constexpr size_t SMEM = 4 * sizeof(Tile) > sizeof(CTile) ? 4 * sizeof(Tile) : sizeof(CTile);
    __shared__ __align__(16) unsigned char smem[SMEM];
    Tile* const As = reinterpret_cast<Tile*>(smem);                     // [2]
    Tile* const Bs = reinterpret_cast<Tile*>(smem + 2 * sizeof(Tile));  // [2]
    const uint32_t MT = (uint32_t)(M + BM2 - 1) / BM2, NT = (uint32_t)(N + BN2 - 1) / BN2;
    uint32_t tile_m, tile_n;
    if (m_inner) {
        const uint32_t q = blockIdx.x / XCDS, NTG = (NT + XCDS - 1) / XCDS, r = q / GM2;
        tile_m = (r / NTG) * GM2 + q % GM2;
        tile_n = (r % NTG) * XCDS + blockIdx.x % XCDS;
    } else {
        tile_n = blockIdx.x % NT;
        tile_m = blockIdx.x / NT;
    }
    if (tile_n >= NT || tile_m >= MT) return;  // padding of the last tile groups (the whole block)
    // the wave index is uniform across the wave: scalar, so that the tile
    // offsets, the head's component and the rows' cell bases stay scalar
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int m0 = (int)tile_m * BM2, n0 = (int)tile_n * BN2;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    const int lr = tid >> 2, lk = (tid & 3) * 8;  // loader: rows lr and lr + 64, k offset
    bf16x8 va[2], vb[2];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int ra = m0 + lr + 64 * h, rb = n0 + lr + 64 * h;
            va[h] = (bf16x8){};
            vb[h] = (bf16x8){};
            if (ra < M) va[h] = *(const bf16x8*)(A + (size_t)ra * lda + k0 + lk);
            if (rb < N) vb[h] = *(const bf16x8*)(B + (size_t)rb * ldb + k0 + lk);
        }
    };
    auto swz = [](int r) { return (r >> 2) & 3; };
    auto stage = [&](int b) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int r = lr + 64 * h, c = ((lk >> 3) ^ swz(r)) << 3;
            *(bf16x8*)&As[b][r][c] = va[h];
            *(bf16x8*)&Bs[b][r][c] = vb[h];
        }
    };
    fetch(0);
    stage(0);
    if (BK < K) fetch(BK);
    const int fr = lane & 15, fq = lane >> 4;
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += BK, buf ^= 1) {
        // stage buf is complete, and every wave is done with stage buf ^ 1
        __syncthreads();
        bf16x8 fa[4], fb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = wm + i * 16 + fr;
            fa[i] = *(const bf16x8*)&As[buf][r][(fq ^ swz(r)) << 3];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = wn + j * 16 + fr;
            fb[j] = *(const bf16x8*)&Bs[buf][r][(fq ^ swz(r)) << 3];
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        if (k0 + BK < K) {
            stage(buf ^ 1);
            if (k0 + 2 * BK < K) fetch(k0 + 2 * BK);
        }
    }
    // Epilogue through LDS, one 16-row slice of each wave's 64 x 64 tile at a
    // time: the slice is written in the MFMA layout and read back row-major,
    // so that consecutive lanes handle consecutive columns (coalesced stores)
    // and the per-output epilogue runs in a loop instead of 64 unrolled copies.
    CTile& Cs = *reinterpret_cast<CTile*>(smem);
    // Head with 64-column-aligned components (cells per block a multiple of
    // 64): a wave's 64 columns share the component, so the 27 far-field
    // samples of each of its rows are the same for the whole wave and come
    // through scalar loads (measured: 5.6 % faster at 512^3 than staging them
    // in LDS), and each lane's 27 interpolation weights stay in registers.
    constexpr bool head_lds = EPI == EPI_HEAD;  // launched only with cells per block % 64 == 0
    float Lx[3], Ly[3], Lz[3];
    uint32_t hc = 0, hcell0 = 0;
    float in_s = 1.0f;
    if (head_lds) {
        const uint32_t n = (uint32_t)(n0 + wn + lane);
        hc = (uint32_t)__builtin_amdgcn_readfirstlane((int)((uint32_t)(n0 + wn) / out.cpb));
        in_s = out.in_scale[hc < 3 ? 0 : 1];
        const uint32_t o = n - hc * out.cpb;
        const uint32_t ox = o % out.bx, oy = (o / out.bx) % out.by, oz = o / (out.bx * out.by);
        lagrange3((float)ox, (float)out.bx, Lx);
        lagrange3((float)oy, (float)out.by, Ly);
        lagrange3((float)oz, (float)out.bz, Lz);
        hcell0 = ox + (oy + oz * out.ny) * out.nx;  // cell offset of the lane within its block
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) Cs[wave][(lane >> 4) * 4 + r][j * 16 + (lane & 15)] = acc[i][j][r];
        __syncthreads();
        if constexpr (head_lds) {
            const int n = n0 + wn + lane;
            const bool train = out.fine != nullptr;
            const bool raw_out = out.raw_out != 0;
            const float rawk = in_s / out.gain;  // network output -> raw source sum
            const float sc = hc < 3 ? out.scale_e : out.scale_b;
            const float kf = hc < 3 ? out.ke : out.kmu;
            const float* stat = (hc < 3 ? out.e_stat : out.b_stat) + (size_t)(hc % 3) * out.ncell;
            float* dyn = (hc < 3 ? out.e_dyn : out.b_dyn) + (size_t)(hc % 3) * out.ncell;
            float* fine = train ? out.fine + (size_t)hc * out.ncell : nullptr;
            // LOD block m -> its first cell: nd is a power of two and cell
            // indices fit 32 bits (at most 2^30 cells)
            const uint32_t lnd = (uint32_t)__builtin_ctz(out.nd), mnd = out.nd - 1u;
            const uint32_t sxb = out.bx, syb = out.by * out.nx, szb = out.bz * out.nx * out.ny;
            // rows in batches of 8: all loads of a batch (flags, static
            // field) are issued before its first store
#pragma unroll
            for (int tb = 0; tb < 16; tb += 8) {
                uint32_t cell[8];
                bool valid[8];
                uint8_t fl[8];
                float st[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int m_raw = m0 + wm + i * 16 + tb + u;
                    valid[u] = m_raw < M && n < N;  // always true for the shapes this path runs
                    // the row's block coordinates are wave-uniform (scalar); only
                    // the lane's cell offset within the block varies
                    const uint32_t m = m_raw < M ? (uint32_t)m_raw : 0u;
                    cell[u] = hcell0 + (uint32_t)__builtin_amdgcn_readfirstlane(
                                           (m & mnd) * sxb + ((m >> lnd) & mnd) * syb + (m >> (2u * lnd)) * szb);
                }
                if (!train) {
#pragma unroll
                    for (int u = 0; u < 8; u++) {  // streamed once: nontemporal, so A stays in L2
                        fl[u] = __builtin_nontemporal_load(out.flags + cell[u]);
                        st[u] = raw_out ? 0.0f : __builtin_nontemporal_load(stat + cell[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int t = tb + u;
                    // separable tri-quadratic interpolation (x, then y, then z)
                    // of the row's 27 far samples, the same for the whole wave:
                    // read through the constant address space (this kernel never
                    // writes them), so that they come as scalar loads although
                    // the kernel stores to other global memory
                    const int m_row = m0 + wm + i * 16 + t;
                    typedef const __attribute__((address_space(4))) float* cfloat_p;
                    const cfloat_p fp =
                        (cfloat_p)(out.far + ((size_t)(m_row < M ? m_row : 0) * 6 + hc) * FAR_PTS);
                    float far = 0.0f;
#pragma unroll
                    for (int zi = 0; zi < 3; zi++) {
                        float fy = 0.0f;
#pragma unroll
                        for (int yi = 0; yi < 3; yi++) {
                            const cfloat_p f = fp + zi * 9 + yi * 3;
                            const float fx = __builtin_fmaf(Lx[2], f[2], __builtin_fmaf(Lx[1], f[1], Lx[0] * f[0]));
                            fy = __builtin_fmaf(Ly[yi], fx, fy);
                        }
                        far = __builtin_fmaf(Lz[zi], fy, far);
                    }
                    const float v = out.alpha * Cs[wave][t][lane];
                    if (!valid[u]) continue;
                    if (train) __builtin_nontemporal_store(v + far * (out.gain / in_s), fine + cell[u]);
                    else if ((fl[u] & 0x1F) != 0x01)
                        __builtin_nontemporal_store(raw_out ? __builtin_fmaf(rawk, v, far) : (st[u] + in_s * sc * v) + kf * far,
                                                    dyn + cell[u]);
                }
            }
        } else {
#pragma unroll 1
            for (int t = 0; t < 16; t++)
                epilogue<EPI>(out, M, N, m0 + wm + i * 16 + t, n0 + wn + lane, Cs[wave][t][lane]);
        }
    }
}

// ---------------------------------------------------------------------------
// Far field: one workgroup per top-level LOD block. Thread t owns row
// r = t % 9 of the block's 3^3 sample lattice (3 points along x that share
// r_y, r_z) and the source chunk c = t / 9 (every 28th source); the sources
// are staged through LDS in tiles and the 28 chunk partials are summed in
// LDS in chunk order (deterministic). Interaction arithmetic: the FAST form
// of the direct kernel (r / |r|^3 with the hardware reciprocal square root).
// Sample k of an axis sits at ((k + 1/2) / 3) b - 1/2 in the block's cell
// coordinates.
// ---------------------------------------------------------------------------
constexpr uint32_t FAR_TILE = 512, FAR_ROWS = 9, FAR_CHUNKS = 28, FAR_THREADS = 256;

__global__ void __launch_bounds__(FAR_THREADS) k_far_samples(FarArgs a) {
    __shared__ float4 s_pq[FAR_TILE];
    __shared__ float4 s_vd[FAR_TILE];
    __shared__ float part[FAR_CHUNKS][FAR_ROWS * 18];
    const uint32_t tid = threadIdx.x, blk = blockIdx.x;
    const uint32_t row = tid % FAR_ROWS, chunk = tid / FAR_ROWS;  // chunk == FAR_CHUNKS: idle lanes
    const uint32_t mask = a.nd - 1u;
    const int bx = (int)(blk & mask), by = (int)((blk >> a.D) & mask), bz = (int)(blk >> (2 * a.D));
    float px[3];
#pragma unroll
    for (int k = 0; k < 3; k++) px[k] = (float)(bx * (int)a.lbx) + ((float)k + 0.5f) * (float)a.lbx / 3.0f - 0.5f;
    const float py = (float)(by * (int)a.lby) + ((float)(row % 3) + 0.5f) * (float)a.lby / 3.0f - 0.5f;
    const float pz = (float)(bz * (int)a.lbz) + ((float)(row / 3) + 0.5f) * (float)a.lbz / 3.0f - 0.5f;
    float acc[3][6];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int c = 0; c < 6; c++) acc[k][c] = 0.0f;
    const uint32_t count = *a.src_count;
    for (uint32_t base = 0; base < count; base += FAR_TILE) {
        const uint32_t m = min(FAR_TILE, count - base);
        __syncthreads();
        for (uint32_t i = tid; i < m; i += FAR_THREADS) {
            s_pq[i] = a.src_pq[base + i];
            s_vd[i] = a.src_vd[base + i];
        }
        __syncthreads();
        if (chunk >= FAR_CHUNKS) continue;
        for (uint32_t s = chunk; s < m; s += FAR_CHUNKS) {
            const float4 vd = s_vd[s];
            const uint32_t d = __float_as_uint(vd.w);
            if (d < a.nblk) {  // in-domain source: skip the 3^3 neighbourhood (own block included)
                const int dx = (int)(d & mask) - bx, dy = (int)((d >> a.D) & mask) - by, dz = (int)(d >> (2 * a.D)) - bz;
                if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) continue;
            }
            const float4 pq = s_pq[s];
            const float ry = py - pq.y, rz = pz - pq.z;
            const float syz = __builtin_fmaf(ry, ry, rz * rz);
            const float cx = __builtin_fmaf(vd.y, rz, -(vd.z * ry));  // (v x r).x
            const float m1 = vd.x * rz, m2 = vd.x * ry;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float rx = px[k] - pq.x;
                const float inv = __builtin_amdgcn_rsqf(__builtin_fmaf(rx, rx, syz));
                const float t = pq.w * (inv * inv * inv);
                acc[k][0] = __builtin_fmaf(t, rx, acc[k][0]);
                acc[k][1] = __builtin_fmaf(t, ry, acc[k][1]);
                acc[k][2] = __builtin_fmaf(t, rz, acc[k][2]);
                acc[k][3] = __builtin_fmaf(t, cx, acc[k][3]);
                acc[k][4] = __builtin_fmaf(t, __builtin_fmaf(vd.z, rx, -m1), acc[k][4]);   // (v x r).y
                acc[k][5] = __builtin_fmaf(t, __builtin_fmaf(-vd.y, rx, m2), acc[k][5]);   // (v x r).z
            }
        }
    }
    if (chunk < FAR_CHUNKS) {
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int c = 0; c < 6; c++) part[chunk][row * 18 + k * 6 + c] = acc[k][c];
    }
    __syncthreads();
    if (tid < FAR_ROWS * 18) {  // tid = row * 18 + k * 6 + c  ->  sample p = k + 3 * row
        float sum = 0.0f;
        for (uint32_t c = 0; c < FAR_CHUNKS; c++) sum += part[c][tid];
        const uint32_t r = tid / 18, k = (tid / 6) % 3, comp = tid % 6;
        a.far[((size_t)blk * 6 + comp) * FAR_PTS + k + 3 * r] = sum;
    }
}

// ---------------------------------------------------------------------------
// Network input: normalised (q, J) of the top-level LODs as bf16 hi + lo
// parts (8 channels), zero for the LOD indices below `lo` that the direct
// field update never reads (sim_kernels.cl:943).
// ---------------------------------------------------------------------------
__global__ void k_input_scale(const float4* __restrict__ qu, uint32_t n_top, float* __restrict__ scale) {
    // scale[0] = max |q|, scale[1] = max |J| (J = q v, the float32 products
    // k_input forms); a zero maximum scales by 1
    __shared__ float sh[2][256];
    float mq = 0.0f, mj = 0.0f;
    for (uint32_t i = threadIdx.x; i < n_top; i += blockDim.x) {
        const float4 e = qu[i];
        mq = fmaxf(mq, fabsf(e.x));
        mj = fmaxf(mj, fmaxf(fabsf(e.x * e.y), fmaxf(fabsf(e.x * e.z), fabsf(e.x * e.w))));
    }
    sh[0][threadIdx.x] = mq;
    sh[1][threadIdx.x] = mj;
    __syncthreads();
    for (uint32_t s = 128; s >= 1; s >>= 1) {
        if (threadIdx.x < s) {
            sh[0][threadIdx.x] = fmaxf(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
            sh[1][threadIdx.x] = fmaxf(sh[1][threadIdx.x], sh[1][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scale[0] = sh[0][0] > 0.0f ? sh[0][0] : 1.0f;
        scale[1] = sh[1][0] > 0.0f ? sh[1][0] : 1.0f;
    }
}

// The charge channel over scale[0], the current channels over scale[1]: E is
// linear in q and B in J, so each output group is scale-free on its own.
__global__ void k_input(const float4* __restrict__ qu, uint32_t n_top, uint32_t lo, const float* __restrict__ scale,
                        __bf16* __restrict__ x) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_top) return;
    const float4 e = qu[v];
    const float inv_q = 1.0f / scale[0], inv_j = 1.0f / scale[1];
    float in[4] = {e.x * inv_q, e.x * e.y * inv_j, e.x * e.z * inv_j, e.x * e.w * inv_j};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float val = v < lo ? 0.0f : in[c];
        const __bf16 hi = f2bf(val);
        x[(size_t)v * 8 + c] = hi;
        x[(size_t)v * 8 + 4 + c] = f2bf(val - bf2f(hi));
    }
}

// im2col of a 3^3 convolution over a g^3 grid (zero padding) of the channel
// concatenation [a (ca channels), b (cb channels)]: row v, column
// k = delta * (ca + cb) + ci, delta = (dx+1) + 3 (dy+1) + 9 (dz+1); columns
// k >= 27 (ca + cb) are zero. transposed != 0 writes the [kpad][ldt] layout.
__global__ void k_im2col(const __bf16* __restrict__ a, int ca, const __bf16* __restrict__ b, int cb, int g, int kpad,
                         __bf16* __restrict__ out, int transposed, int ldt) {
    const int cin = ca + cb;
    const size_t total = (size_t)g * g * g * kpad;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(idx / kpad), k = (int)(idx % kpad);
        __bf16 val = f2bf(0.0f);
        if (k < 27 * cin) {
            const int d = k / cin, ci = k % cin;
            const int x = v % g + d % 3 - 1, y = (v / g) % g + (d / 3) % 3 - 1, z = v / (g * g) + d / 9 - 1;
            if (x >= 0 && y >= 0 && z >= 0 && x < g && y < g && z < g) {
                const int w = x + g * (y + g * z);
                val = ci < ca ? a[(size_t)w * ca + ci] : b[(size_t)w * cb + (ci - ca)];
            }
        }
        if (transposed) out[(size_t)k * ldt + v] = val;
        else out[idx] = val;
    }
}

// col2im (adjoint of im2col): d_in[w][ci] = sum over delta of dA[w - delta][delta*cin + ci],
// written split into the two concatenated inputs (fp32 gradients).
__global__ void k_col2im(const float* __restrict__ dA, int g, int kpad, int ca, int cb, float* __restrict__ da,
                         float* __restrict__ db) {
    const int cin = ca + cb;
    const size_t total = (size_t)g * g * g * cin;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int w = (int)(idx / cin), ci = (int)(idx % cin);
        const int wx = w % g, wy = (w / g) % g, wz = w / (g * g);
        float s = 0.0f;
        for (int d = 0; d < 27; d++) {
            const int x = wx - (d % 3 - 1), y = wy - ((d / 3) % 3 - 1), z = wz - (d / 9 - 1);
            if (x < 0 || y < 0 || z < 0 || x >= g || y >= g || z >= g) continue;
            s += dA[(size_t)(x + g * (y + g * z)) * kpad + d * cin + ci];
        }
        if (ci < ca) {
            if (da) da[(size_t)w * ca + ci] = s;
        } else if (db) {
            db[(size_t)w * cb + (ci - ca)] = s;
        }
    }
}

// 2x2x2 average pooling (g -> g/2) and nearest upsampling (g/2 -> g), bf16
// activations; the backward versions work on fp32 gradients.
__global__ void k_pool(const __bf16* __restrict__ in, int g, int c, __bf16* __restrict__ out) {
    const int h = g / 2;
    const size_t total = (size_t)h * h * h * c;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(idx / c), ch = (int)(idx % c);
        const int x = 2 * (v % h), y = 2 * ((v / h) % h), z = 2 * (v / (h * h));
        float s = 0.0f;
        for (int d = 0; d < 8; d++)
            s += bf2f(in[(size_t)((x + (d & 1)) + g * ((y + ((d >> 1) & 1)) + g * (z + (d >> 2)))) * c + ch]);
        out[idx] = f2bf(0.125f * s);
    }
}
__global__ void k_pool_bwd(const float* __restrict__ dout, int g, int c, float* __restrict__ din, int accumulate) {
    const int h = g / 2;
    const size_t total = (size_t)g * g * g * c;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(idx / c), ch = (int)(idx % c);
        const int x = (v % g) / 2, y = ((v / g) % g) / 2, z = (v / (g * g)) / 2;
        const float val = 0.125f * dout[(size_t)(x + h * (y + h * z)) * c + ch];
        din[idx] = accumulate ? din[idx] + val : val;
    }
}
__global__ void k_up(const __bf16* __restrict__ in, int h, int c, __bf16* __restrict__ out) {
    const int g = 2 * h;
    const size_t total = (size_t)g * g * g * c;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(idx / c), ch = (int)(idx % c);
        const int x = (v % g) / 2, y = ((v / g) % g) / 2, z = (v / (g * g)) / 2;
        out[idx] = in[(size_t)(x + h * (y + h * z)) * c + ch];
    }
}
__global__ void k_up_bwd(const float* __restrict__ dout, int h, int c, float* __restrict__ din) {
    const int g = 2 * h;
    const size_t total = (size_t)h * h * h * c;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(idx / c), ch = (int)(idx % c);
        const int x = 2 * (v % h), y = 2 * ((v / h) % h), z = 2 * (v / (h * h));
        float s = 0.0f;
        for (int d = 0; d < 8; d++)
            s += dout[(size_t)((x + (d & 1)) + g * ((y + ((d >> 1) & 1)) + g * (z + (d >> 2)))) * c + ch];
        din[idx] = s;
    }
}

// fp32 [rows][cols] -> bf16 [rows][ld] (zero in padding columns), or transposed
// to [ld][ldt] (ldt >= rows; columns rows..ldt-1 are left untouched).
__global__ void k_to_bf16(const float* __restrict__ in, int rows, int cols, int ld, __bf16* __restrict__ out,
                          int transposed, int ldt) {
    const size_t total = (size_t)rows * ld;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(idx / ld), cc = (int)(idx % ld);
        const __bf16 v = f2bf(cc < cols ? in[(size_t)r * cols + cc] : 0.0f);
        if (transposed) {
            out[(size_t)cc * ldt + r] = v;  // out is [ld][ldt], ldt >= rows
        } else {
            out[idx] = v;
        }
    }
}

// Loss of the head output y (network units, fp32 [6][N]) against the direct
// field's raw source sums (E then B, [6][N]) in network units,
// t = raw * gain / s (s = the charge scale for E, the current scale for B),
// over fluid cells: per-block partial sums of (y - t)^2
// and t^2 for the E and B groups.
__global__ void __launch_bounds__(256) k_loss_sums(const float* __restrict__ y, const float* __restrict__ target,
                                                   const uint8_t* __restrict__ flags, uint64_t ncell,
                                                   const float* __restrict__ in_scale, float gain,
                                                   double* __restrict__ partial) {
    __shared__ double sh[4][256];
    double s[4] = {0, 0, 0, 0};
    const float tge = gain / in_scale[0], tgb = gain / in_scale[1];
    for (uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; n < ncell; n += (uint64_t)gridDim.x * blockDim.x) {
        if ((flags[n] & 0x1F) == 0x01) continue;
        for (int c = 0; c < 6; c++) {
            const size_t i = (size_t)c * ncell + n;
            const float t = target[i] * (c < 3 ? tge : tgb);
            const float d = y[i] - t;
            s[(c < 3 ? 0 : 2) + 0] += (double)d * d;
            s[(c < 3 ? 0 : 2) + 1] += (double)t * t;
        }
    }
    for (int k = 0; k < 4; k++) sh[k][threadIdx.x] = s[k];
    __syncthreads();
    for (uint32_t st = 128; st >= 1; st >>= 1) {
        if (threadIdx.x < st)
            for (int k = 0; k < 4; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) partial[(size_t)blockIdx.x * 4 + k] = sh[k][0];
}

// dL/dy for L = sum_E (y-t)^2 / sum_E t^2 + sum_B (y-t)^2 / sum_B t^2, written as
// bf16 in the two GEMM layouts dY[block][n] and dYt[n][block] (n = c*cpb + o);
// the padding columns of both stay zero.
__global__ void k_loss_grad(const float* __restrict__ y, const float* __restrict__ target,
                            const uint8_t* __restrict__ flags, HeadGeom hg, const float* __restrict__ in_scale,
                            float gain, const float* __restrict__ norms, __bf16* __restrict__ dY,
                            __bf16* __restrict__ dYt) {
    const uint64_t ncell = hg.ncell;
    const float tge = gain / in_scale[0], tgb = gain / in_scale[1];
    const float ge = norms[0] > 0.0f ? 2.0f / norms[0] : 0.0f, gb = norms[1] > 0.0f ? 2.0f / norms[1] : 0.0f;
    for (uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; n < ncell; n += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t z = (uint32_t)(n / ((uint64_t)hg.nx * hg.ny)), r = (uint32_t)(n % ((uint64_t)hg.nx * hg.ny));
        const uint32_t yy = r / hg.nx, x = r % hg.nx;
        const uint32_t blk = x / hg.bx + (yy / hg.by + (z / hg.bz) * hg.nd) * hg.nd;
        const uint32_t o = (x % hg.bx) + ((yy % hg.by) + (z % hg.bz) * hg.by) * hg.bx;
        const bool fluid = (flags[n] & 0x1F) != 0x01;
        for (int c = 0; c < 6; c++) {
            const size_t i = (size_t)c * ncell + n;
            const float g = fluid ? (c < 3 ? ge : gb) * (y[i] - target[i] * (c < 3 ? tge : tgb)) : 0.0f;
            const uint32_t col = c * hg.cpb + o;
            dY[(size_t)blk * hg.ld_dy + col] = f2bf(g);
            dYt[(size_t)col * hg.ld_dyt + blk] = f2bf(g);
        }
    }
}

// Adam on fp32 master weights; refreshes the bf16 copy ([rows][ld] with the
// master's [rows][cols] data and zero padding).
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       size_t n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

// Adam for the head matrix [rows][kpad]: columns of level-0 features
// (k % hcin < c0) step with lr_feat, columns of the direct source channels
// with lr_src.
__global__ void k_adam_head(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, size_t n, int kpad, int hcin, int c0, float lr_feat, float lr_src,
                            float b1, float b2, float eps, float bc1, float bc2) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(i % (size_t)kpad);
        const float lr = (k % hcin) < c0 ? lr_feat : lr_src;
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

// ---------------------------------------------------------------------------
static inline dim3 grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 65535) b = 65535;
    if (b == 0) b = 1;
    return dim3((uint32_t)b);
}

// Sum of the split-K partials [splits][M][N] in split order, then the epilogue.
template <int EPI>
__global__ void __launch_bounds__(256) k_splitk_reduce(GemmOut out, int M, int N, int splits) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * N) return;
    const float* ws = out.splitk_ws;
    float acc = ws[idx];
    for (int z = 1; z < splits; z++) acc += ws[(size_t)z * M * N + idx];
    epilogue<EPI>(out, M, N, idx / N, idx % N, acc);
}

hipError_t gemm_nt(int epi, const __bf16* A, const __bf16* B, int M, int N, int K, int lda, int ldb, const GemmOut& o,
                   hipStream_t s) {
    if (K % BK != 0 || M <= 0 || N <= 0) return hipErrorInvalidValue;
    if (M >= BM2 && N >= BN2 && (epi != EPI_HEAD || (o.cpb & 63u) == 0)) {
        const uint32_t MT = (uint32_t)(M + BM2 - 1) / BM2, NT = (uint32_t)(N + BN2 - 1) / BN2;
        const int mi = NT > MT;
        const dim3 g(mi ? XCDS * GM2 * ((NT + XCDS - 1) / XCDS) * ((MT + GM2 - 1) / GM2) : MT * NT);
        if (epi == EPI_F32) hipLaunchKernelGGL(k_gemm_nt_128<EPI_F32>, g, dim3(256), 0, s, A, B, M, N, K, lda, ldb, o, mi);
        else if (epi == EPI_BF16) hipLaunchKernelGGL(k_gemm_nt_128<EPI_BF16>, g, dim3(256), 0, s, A, B, M, N, K, lda, ldb, o, mi);
        else hipLaunchKernelGGL(k_gemm_nt_128<EPI_HEAD>, g, dim3(256), 0, s, A, B, M, N, K, lda, ldb, o, mi);
        return hipGetLastError();
    }
    const dim3 g((N + BN - 1) / BN, (M + BM - 1) / BM);
    // Few tiles and a long K (the U-Net's deep layers: 1-64 tiles, 7-95 K
    // slices): split K so that at least ~256 blocks run, each over a run of
    // K slices, into fp32 partials that a second kernel sums in a fixed order.
    const int nk = K / BK, tiles = (int)(g.x * g.y);
    int splits = 1;
    if (o.splitk_ws != nullptr && (epi == EPI_F32 || epi == EPI_BF16) && tiles < 256 && nk >= 4) {
        splits = std::min(std::min(nk / 2, 32), (256 + tiles - 1) / tiles);
        while (splits > 1 && (size_t)splits * M * N > o.splitk_cap) splits--;
        const int per = (nk + splits - 1) / splits;
        splits = (nk + per - 1) / per;  // no empty run
    }
    if (splits > 1) {
        GemmOut po = o;
        po.f32 = o.splitk_ws;
        hipLaunchKernelGGL(k_gemm_nt<EPI_PARTIAL>, dim3(g.x, g.y, (uint32_t)splits), dim3(256), 0, s, A, B, M, N, K, lda,
                           ldb, po);
        const dim3 gr((uint32_t)((M * N + 255) / 256));
        if (epi == EPI_F32) hipLaunchKernelGGL(k_splitk_reduce<EPI_F32>, gr, dim3(256), 0, s, o, M, N, splits);
        else hipLaunchKernelGGL(k_splitk_reduce<EPI_BF16>, gr, dim3(256), 0, s, o, M, N, splits);
        return hipGetLastError();
    }
    if (epi == EPI_F32) hipLaunchKernelGGL(k_gemm_nt<EPI_F32>, g, dim3(256), 0, s, A, B, M, N, K, lda, ldb, o);
    else if (epi == EPI_BF16) hipLaunchKernelGGL(k_gemm_nt<EPI_BF16>, g, dim3(256), 0, s, A, B, M, N, K, lda, ldb, o);
    else hipLaunchKernelGGL(k_gemm_nt<EPI_HEAD>, g, dim3(256), 0, s, A, B, M, N, K, lda, ldb, o);
    return hipGetLastError();
}
hipError_t far_samples(const FarArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_far_samples, dim3(a.nblk), dim3(FAR_THREADS), 0, s, a);
    return hipGetLastError();
}
hipError_t input(const float4* qu, uint32_t n_top, uint32_t lo, float* scale, __bf16* x, hipStream_t s) {
    hipLaunchKernelGGL(k_input_scale, dim3(1), dim3(256), 0, s, qu, n_top, scale);
    hipLaunchKernelGGL(k_input, dim3((n_top + 255) / 256), dim3(256), 0, s, qu, n_top, lo, scale, x);
    return hipGetLastError();
}
hipError_t im2col(const __bf16* a, int ca, const __bf16* b, int cb, int g, int kpad, __bf16* out, int transposed,
                  int ldt, hipStream_t s) {
    hipLaunchKernelGGL(k_im2col, grid_for((size_t)g * g * g * kpad), dim3(256), 0, s, a, ca, b, cb, g, kpad, out,
                       transposed, ldt);
    return hipGetLastError();
}
hipError_t col2im(const float* dA, int g, int kpad, int ca, int cb, float* da, float* db, hipStream_t s) {
    hipLaunchKernelGGL(k_col2im, grid_for((size_t)g * g * g * (ca + cb)), dim3(256), 0, s, dA, g, kpad, ca, cb, da, db);
    return hipGetLastError();
}
hipError_t pool(const __bf16* in, int g, int c, __bf16* out, hipStream_t s) {
    hipLaunchKernelGGL(k_pool, grid_for((size_t)(g / 2) * (g / 2) * (g / 2) * c), dim3(256), 0, s, in, g, c, out);
    return hipGetLastError();
}
hipError_t pool_bwd(const float* dout, int g, int c, float* din, int accumulate, hipStream_t s) {
    hipLaunchKernelGGL(k_pool_bwd, grid_for((size_t)g * g * g * c), dim3(256), 0, s, dout, g, c, din, accumulate);
    return hipGetLastError();
}
hipError_t up(const __bf16* in, int h, int c, __bf16* out, hipStream_t s) {
    hipLaunchKernelGGL(k_up, grid_for((size_t)8 * h * h * h * c), dim3(256), 0, s, in, h, c, out);
    return hipGetLastError();
}
hipError_t up_bwd(const float* dout, int h, int c, float* din, hipStream_t s) {
    hipLaunchKernelGGL(k_up_bwd, grid_for((size_t)h * h * h * c), dim3(256), 0, s, dout, h, c, din);
    return hipGetLastError();
}
hipError_t to_bf16(const float* in, int rows, int cols, int ld, __bf16* out, int transposed, int ldt, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, grid_for((size_t)rows * ld), dim3(256), 0, s, in, rows, cols, ld, out, transposed, ldt);
    return hipGetLastError();
}
hipError_t loss_sums(const float* y, const float* target, const uint8_t* flags, uint64_t ncell, const float* in_scale,
                     float gain, double* partial, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_loss_sums, dim3(nblocks), dim3(256), 0, s, y, target, flags, ncell, in_scale, gain, partial);
    return hipGetLastError();
}
hipError_t loss_grad(const float* y, const float* target, const uint8_t* flags, const HeadGeom& hg,
                     const float* in_scale, float gain, const float* norms, __bf16* dY, __bf16* dYt, hipStream_t s) {
    hipLaunchKernelGGL(k_loss_grad, grid_for(hg.ncell), dim3(256), 0, s, y, target, flags, hg, in_scale, gain, norms,
                       dY, dYt);
    return hipGetLastError();
}
hipError_t adam_head(float* w, const float* g, float* m, float* v, size_t n, int kpad, int hcin, int c0, float lr_feat,
                     float lr_src, float b1, float b2, float eps, float bc1, float bc2, hipStream_t s) {
    hipLaunchKernelGGL(k_adam_head, grid_for(n), dim3(256), 0, s, w, g, m, v, n, kpad, hcin, c0, lr_feat, lr_src, b1,
                       b2, eps, bc1, bc2);
    return hipGetLastError();
}
hipError_t adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2, float eps,
                float bc1, float bc2, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, grid_for(n), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, bc1, bc2);
    return hipGetLastError();
}

}  // namespace sur
}  // namespace ion
