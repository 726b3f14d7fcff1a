// Device side of mesh voxelisation and the static-field precompute.
//
// Reference: voxelize_mesh, psi_from_mesh, static_b_from_mesh and
// static_e_from_mesh (src/kernels/sim_kernels.cl:1150-1300) with their host
// drivers (src/mesh.rs:256-343, src/lbm/mod.rs:284-314,
// src/lbm/domain.rs:551-567).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ion {
namespace mesh {

// Flag bits of sim_kernels.cl:45-52.
constexpr uint8_t TYPE_S = 0x01, TYPE_C = 0x04, TYPE_F = 0x08, TYPE_M = 0x10;
// Runtime TYPE_BO (src/lbm/domain.rs:828): the bits a voxelised flag replaces.
constexpr uint8_t TYPE_BO = 0x1F;
// Ray-mesh intersections kept per column (sim_kernels.cl:1174).
constexpr uint32_t MAX_HITS = 64;

// One voxelisation pass: rays along `direction` through every column of the
// lattice face of that direction, against T triangles stored as
// tri[9 t .. 9 t + 8] = p0, p1, p2 (x, y, z each).
struct VoxArgs {
    const float* tri;
    uint32_t T;
    uint32_t direction;  // 0: x, 1: y, 2: z
    uint32_t A;          // columns: ny nz, nz nx or nx ny
    uint32_t nx, ny, nz;
    uint64_t N;
    uint8_t* flags;
    float* b_dyn;        // null without MAGNETO_HYDRO
    uint32_t flag;       // the flag written into voxelised cells
    float mx, my, mz;    // magnetisation per cell (magnets) or charge in mx (charged)
    float x0, y0, z0, x1, y1, z1;  // mesh bounds widened by 2 cells
};

// Cells selected as field sources, in ascending cell order: position (cell
// coordinates plus `offset`) and value (magnetisation, or charge in .x).
struct Sources {
    float4* pos;
    float4* val;
    // count[0]: number of sources; count[1]: nonzero if a magnet's
    // magnetisation lies outside the range of the specialised division
    // (the scatter pass sets it; cleared by the caller before)
    uint32_t* count;
};

hipError_t launch_voxelize(const VoxArgs& a, hipStream_t s);

// Source selection in two passes, the host reading the total in between:
// mode 0: magnets (flags & TYPE_M), value = (M[i], M[N+i], M[2N+i]), offset 1
// mode 1: charges (flags & (TYPE_F | TYPE_C)), value.x = C[i], offset 0
// block_counts: select_blocks(N) + 1 uint32; the count pass leaves the
// exclusive block offsets there and the total in *total.
uint32_t select_blocks(uint64_t N);
hipError_t launch_select_count(int mode, const uint8_t* flags, uint64_t N, uint32_t* block_counts, uint32_t* total,
                               hipStream_t s);
hipError_t launch_select_scatter(int mode, const uint8_t* flags, const float* m, uint64_t N, uint32_t nx, uint32_t ny,
                                 const uint32_t* block_counts, Sources out, hipStream_t s);

// Exhaustive / sampled device checks of the specialised correctly rounded
// square root and division of k_psi against the compiler's general ones.
hipError_t run_psi_math_check(uint64_t* sqrt_mismatch, uint64_t* div_mismatch, uint64_t* div_samples);

// psi over the lattice padded by one cell on every side
// ((nx+2)(ny+2)(nz+2) values), from the magnet sources.
hipError_t launch_psi(const Sources& src, float* psi, uint32_t nx, uint32_t ny, uint32_t nz, hipStream_t s);
// B += -kmu0 grad(psi) (central differences) for cells without the solid bit.
hipError_t launch_static_b(const uint8_t* flags, float* b, const float* psi, uint32_t nx, uint32_t ny, uint32_t nz,
                           float kmu0, hipStream_t s);
// E += ke * sum over charged cells of (c - c_i) q_i / |c - c_i|^3, for cells
// without the solid bit.
hipError_t launch_static_e(const Sources& src, const uint8_t* flags, float* e, uint32_t nx, uint32_t ny, uint32_t nz,
                           float ke, hipStream_t s);

}  // namespace mesh
}  // namespace ion
