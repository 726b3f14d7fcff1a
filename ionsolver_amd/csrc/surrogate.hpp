// Kernels of the learned field surrogate (surrogate.hip) and its host-side
// network (surrogate_host.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ion {
namespace sur {

enum { EPI_F32 = 0, EPI_BF16 = 1, EPI_HEAD = 2, EPI_PARTIAL = 3 };

// Output of a GEMM: fp32 / bf16 matrix with leading dimension ldc, or the
// head epilogue that scatters block x (component, offset) onto the cells.
struct GemmOut {
    float* f32;
    __bf16* bf16;
    int ldc;
    int accumulate;
    float alpha;
    // head epilogue
    float* fine;              // training: normalised output [6][ncell]
    float* e_dyn;
    float* b_dyn;
    const float* e_stat;
    const float* b_stat;
    const uint8_t* flags;
    const float* in_scale;    // input normalisation (device [2]: charge scale for E, current scale for B)
    float scale_e, scale_b;   // KE / gain, KMU / gain (gain: network output units per raw field unit)
    float ke, kmu, gain;
    const float* far;         // far-field samples [block][6][27], interpolated in the epilogue
    int raw_out;              // write the raw source sums of the fluid cells into e_dyn / b_dyn (for the judge)
    uint64_t ncell;
    uint32_t nx, ny, nz, nd, bx, by, bz, cpb;
    // split K (64 x 64 tiles, EPI_F32 / EPI_BF16): fp32 partial sums
    // [splits][M][N] of up to splitk_cap floats, summed in split order by a
    // second kernel that applies the epilogue (null: no split)
    float* splitk_ws;
    size_t splitk_cap;
};

struct HeadGeom {
    uint64_t ncell;
    uint32_t nx, ny, nz, nd, bx, by, bz, cpb;
    uint32_t ld_dy, ld_dyt;  // leading dimensions of dY [block][ld_dy] and dYt [n][ld_dyt]
};

// Far field of a top-level LOD block: the sources outside its 3^3 block
// neighbourhood (the ones the head's near-field stencil does not see),
// sampled at a 3^3 lattice of points inside the block.
constexpr int FAR_PTS = 27;
struct FarArgs {
    const float4* src_pq;      // compacted LOD sources (position, charge)
    const float4* src_vd;      // (velocity, LOD index bits)
    const uint32_t* src_count;
    float* far;                // [block][6][27]: E then B raw sums, the 27 points of a component together
    uint32_t nd, D, nblk;
    uint32_t lbx, lby, lbz;    // cells per block side
};
hipError_t far_samples(const FarArgs& a, hipStream_t s);
hipError_t gemm_nt(int epi, const __bf16* A, const __bf16* B, int M, int N, int K, int lda, int ldb, const GemmOut& o,
                   hipStream_t s);
hipError_t input(const float4* qu, uint32_t n_top, uint32_t lo, float* scale, __bf16* x, hipStream_t s);
hipError_t im2col(const __bf16* a, int ca, const __bf16* b, int cb, int g, int kpad, __bf16* out, int transposed,
                  int ldt, hipStream_t s);
hipError_t col2im(const float* dA, int g, int kpad, int ca, int cb, float* da, float* db, hipStream_t s);
hipError_t pool(const __bf16* in, int g, int c, __bf16* out, hipStream_t s);
hipError_t pool_bwd(const float* dout, int g, int c, float* din, int accumulate, hipStream_t s);
hipError_t up(const __bf16* in, int h, int c, __bf16* out, hipStream_t s);
hipError_t up_bwd(const float* dout, int h, int c, float* din, hipStream_t s);
hipError_t to_bf16(const float* in, int rows, int cols, int ld, __bf16* out, int transposed, int ldt, hipStream_t s);
hipError_t loss_sums(const float* y, const float* target, const uint8_t* flags, uint64_t ncell, const float* in_scale,
                     float gain, double* partial, int nblocks, hipStream_t s);
hipError_t loss_grad(const float* y, const float* target, const uint8_t* flags, const HeadGeom& hg,
                     const float* in_scale, float gain, const float* norms, __bf16* dY, __bf16* dYt, hipStream_t s);
hipError_t adam_head(float* w, const float* g, float* m, float* v, size_t n, int kpad, int hcin, int c0, float lr_feat,
                     float lr_src, float b1, float b2, float eps, float bc1, float bc2, hipStream_t s);
hipError_t adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2, float eps,
                float bc1, float bc2, hipStream_t s);

}  // namespace sur
}  // namespace ion
