// Library-internal view of a handle for the host modules that live outside
// lbm_host.cpp (mesh import, voxelisation, static fields).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ionsolver.h"

namespace ion {

struct MeshSet;                    // meshes imported into a handle (mesh_host.cpp)
void mesh_set_destroy(MeshSet*);   // called when the handle is destroyed

struct HandleView {
    hipStream_t stream;
    uint32_t nx, ny, nz;
    uint64_t N;
    bool mhd, f64;
    const ion_config* cfg;
    const ion_kernel_constants* kc;
    uint8_t* flags;
    float *e_stat, *b_stat, *e_dyn, *b_dyn;
    float* e_var;  // SUBGRID_ECR only, else null
    MeshSet** meshes;
};

}  // namespace ion

extern "C" {
__attribute__((visibility("hidden"))) int ion_internal_view(ion_handle* h, ion::HandleView* v);
__attribute__((visibility("hidden"))) int ion_internal_set_error(int code, const char* msg);
}
