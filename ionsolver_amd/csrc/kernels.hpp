// Kernel argument blocks and launchers shared by the host runtime
// (lbm_host.cpp) and the gfx950 kernels (kernels_lbm.hip, kernels_field.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ion {

enum StoreKind { STORE_FP16S = 0, STORE_FP16C = 1, STORE_FP32 = 2, STORE_FP64 = 3 };

// Arguments of initialize / stream_collide / update_fields. Buffers are the
// reference's device buffers (src/lbm/domain.rs:151-211), all SoA:
// scalar fields [N], vector fields [3][N], DDFs [q][N].
struct LbmArgs {
    void* fi;
    float* rho;
    float* u;
    const uint8_t* flags;
    const float* F;       // force field (FORCE_FIELD), else null
    float* e_stat;
    float* b_stat;
    float* e_dyn;
    float* b_dyn;
    void* fqi;            // D3Q7 charge DDFs
    void* ei;             // electron-gas DDFs
    float* q;             // cell charge
    float4* lod_seg;      // level-1 LOD row-segment sums [N / lod_bx]
    uint32_t nx, ny, nz;
    uint32_t pad0;
    uint64_t N;
    uint64_t N_store;     // == N; a separate argument so that the compiler computes the
                          // store-side slot bases at the stores instead of keeping the
                          // load-side ones live in scalar registers across the cell
    uint32_t odd;         // t % 2 (esoteric-pull layout)
    uint32_t eqb, vforce, ffield, upd;  // extension switches
    uint32_t qin;         // charge inlet: TYPE_E cells hold their charge (build extension)
    void* eti;            // SUBGRID_ECR: electron temperature DDFs (D3Q7)
    float* et;            // SUBGRID_ECR: electron temperature field
    float* e_var;         // SUBGRID_ECR: oscillating electric field
    float kkbme, keabs;   // DEF_KKBME, DEF_KEABS
    uint32_t lod_bx;      // LOD block side along x (cells)
    float fx, fy, fz;     // volume force
    float w;              // DEF_W
    float c;              // DEF_C
    float kkge, wq;       // DEF_KKGE, DEF_WQ
    float ils;            // 1 / cells per LOD block
    double ils_d;         // the same quotient in double (FP64 instantiation)
    float ei_rho0;        // initialize: electron-gas density of non-solid cells (build extension; 0 = reference)
};

// Arguments of the E/B field update.
struct FieldArgs {
    const float* e_stat;
    const float* b_stat;
    float* e_dyn;
    float* b_dyn;
    const uint8_t* flags;
    const float* q;
    const float* u;
    const float4* src_pq;   // LOD sources: centre x,y,z and charge
    const float4* src_vd;   // LOD sources: mean velocity and LOD index (bits)
    const uint32_t* src_count;
    uint32_t nx, ny, nz;
    uint32_t D;
    uint64_t N;
    uint32_t dsx, dsy, dsz; // near-field block side
    uint32_t near;          // near-field block has more than one cell
    float ke, kmu;
    float* raw;             // if set: the raw source sums (E then B, 6 x N) of fluid cells instead of E_dyn / B_dyn
    uint32_t f64;           // FP64 instantiation: every field and source array holds doubles
    // judge: E_dyn / B_dyn hold the surrogate's raw sums for the fluid cells on
    // entry; the kernel adds (surrogate - direct)^2 and direct^2 of the E and
    // the B components into judge[block][4] before writing its own fields
    double* judge;
};
// Number of blocks (judge partials) the field kernel launches for a.
uint32_t field_blocks(const FieldArgs& a, bool fast);
// Sums judge partials [nblocks][4] in a fixed order into out[4] (E err, E norm, B err, B norm).
hipError_t launch_judge_reduce(const double* partial, uint32_t nblocks, double* out, hipStream_t s);

// kernels_lbm.hip
// ecr: the SUBGRID_ECR variants (D3Q19, MHD, FP32/FP16S/FP16C)
hipError_t launch_initialize(int q, int store, bool mhd, const LbmArgs& a, hipStream_t s, bool ecr = false);
hipError_t launch_stream_collide(int q, int store, bool mhd, bool trt, bool lod, uint32_t block,
                                 const LbmArgs& a, hipStream_t s, bool ecr = false);
hipError_t launch_update_fields(int q, int store, const LbmArgs& a, hipStream_t s);
// Level-2 LOD tree: one block per top-level LOD.
hipError_t launch_lod_reduce(const void* seg, void* qu_lod, bool f64, uint32_t nx, uint32_t ny, uint32_t nz,
                             uint32_t D, hipStream_t s);
// fill n floats (or doubles when f64) with v
hipError_t launch_fill(void* p, bool f64, uint64_t n, double v, hipStream_t s);
hipError_t launch_codec(int store, int dir, const void* in, void* out, uint64_t n, hipStream_t s);

// kernels_field.hip
hipError_t launch_prep_sources(const void* qu_lod, bool f64, uint32_t lo, uint32_t hi, uint32_t nx, uint32_t ny,
                               uint32_t nz, uint32_t D, void* src_pq, void* src_vd, uint32_t* count, hipStream_t s);
hipError_t launch_update_e_b(const FieldArgs& a, bool fast, hipStream_t s);
// E_dyn = E_stat + KE raw_E, B_dyn = B_stat + KMU raw_B for fluid cells (the
// direct kernel's own epilogue, applied to a.raw).
hipError_t launch_combine_fields(const FieldArgs& a, hipStream_t s);
hipError_t launch_sum_f64(const void* x, bool f64, const uint8_t* flags, uint64_t n, double* partial,
                          uint32_t nblocks, hipStream_t s);

}  // namespace ion
