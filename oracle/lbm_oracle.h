/*
 * CPU oracle for the IonSolver extended-LBM MHD time step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else. The product library (libionsolver_hip.so) never links it.
 *
 * It restates the semantics of the reference's OpenCL kernels
 * (src/kernels/sim_kernels.cl) and of the host step sequencing
 * (src/lbm/mod.rs:214-272) in portable C++, templated on the DDF storage
 * format (FP32 / FP16S / FP16C, plus FP64 for the double-precision oracle)
 * and on the arithmetic type (float, or double for the FP64 oracle).
 * Where a reference order of operations is fixed by the source text, the
 * oracle follows it literally (explicit fma() only where the source writes
 * fma(); everything else is a separate, correctly rounded IEEE operation).
 *
 * The one place the reference has no defined order, the float atomics that
 * build the level-of-detail (LOD) charge/velocity sums
 * (sim_kernels.cl:666-677), is replaced by a fixed pairwise-tree order:
 *   level 1: for each row segment (y, z, lod-x-block) the bx cell values,
 *            padded with +0 to the next power of two P, are summed by
 *            for (s = P/2; s >= 1; s /= 2) a[i] = a[i] + a[i+s] (i < s);
 *   level 2: for each LOD block the by*bz row-segment sums, indexed
 *            j = (y - Y*by) + (z - Z*bz)*by, are summed by the same tree.
 * Solid cells contribute +0. The GPU kernels use the same order, so the two
 * agree exactly.
 */
#ifndef IONSOLVER_LBM_ORACLE_H
#define IONSOLVER_LBM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Extension bits (LbmConfig ext_* flags, src/lbm/mod.rs:73-84, plus
 * UPDATE_FIELDS which the reference derives from graphics_active,
 * src/lbm/domain.rs:856). */
enum {
    ORC_EXT_EQUILIBRIUM_BOUNDARIES = 1 << 0,
    ORC_EXT_VOLUME_FORCE = 1 << 1,
    ORC_EXT_FORCE_FIELD = 1 << 2,
    ORC_EXT_MAGNETO_HYDRO = 1 << 3,
    ORC_EXT_UPDATE_FIELDS = 1 << 4,
    /* Charge inlet, a build-defined extension with no reference counterpart:
     * TYPE_E cells (with EQUILIBRIUM_BOUNDARIES and MAGNETO_HYDRO) take the
     * charge density from q[n], reset their D3Q7 DDFs to a_eq(q[n], u) with
     * no collision and leave q[n] unwritten. */
    ORC_EXT_CHARGE_INLET = 1 << 5,
    /* SUBGRID_ECR (sim_kernels.cl:557-629, 827-831): electron temperature
     * DDFs, ECR heating by E_var, gyration drift, ionisation. */
    ORC_EXT_SUBGRID_ECR = 1 << 6,
};

/* Storage formats: discriminants of FloatType (src/lbm/types.rs:75-83), plus
 * FP64 for the double-precision oracle. */
enum { ORC_FP16S = 0, ORC_FP16C = 1, ORC_FP32 = 2, ORC_FP64 = 3 };
/* Velocity sets: discriminants of VelocitySet (src/lbm/types.rs:16-26). */
enum { ORC_D2Q9 = 0, ORC_D3Q15 = 1, ORC_D3Q19 = 2, ORC_D3Q27 = 3 };

/* Kernel-level parameters, i.e. the values the reference bakes into the
 * OpenCL program as #defines (src/lbm/domain.rs:786-857). Physical
 * constants are the f32 values the host derives; the FP64 oracle promotes
 * them to double. */
typedef struct orc_params {
    int32_t velocity_set;     /* ORC_D2Q9 .. ORC_D3Q19 */
    int32_t relaxation_time;  /* 0 = SRT, 1 = TRT */
    int32_t float_type;       /* ORC_FP16S .. ORC_FP64 */
    uint32_t nx, ny, nz;
    uint32_t lod_depth;       /* DEF_LOD_DEPTH */
    uint32_t ext;             /* ORC_EXT_* bits */
    float w;                  /* DEF_W = 1/(3 nu + 1/2) */
    float fx, fy, fz;         /* volume force kernel arguments */
    float ke, kmu, kkge, wq;  /* DEF_KE, DEF_KMU, DEF_KKGE, DEF_WQ */
    float kkbme, keabs;       /* DEF_KKBME, DEF_KEABS (SUBGRID_ECR) */
    float ei_rho0;            /* initial electron-gas density of non-solid cells
                                 (build extension; 0 = sim_kernels.cl:823-825) */
} orc_params;

/* Buffer ids, one per reference device buffer (src/lbm/domain.rs:151-211). */
enum {
    ORC_BUF_FI = 0, ORC_BUF_RHO, ORC_BUF_U, ORC_BUF_FLAGS, ORC_BUF_F,
    ORC_BUF_E_STAT, ORC_BUF_B_STAT, ORC_BUF_E_DYN, ORC_BUF_B_DYN,
    ORC_BUF_FQI, ORC_BUF_Q, ORC_BUF_QU_LOD, ORC_BUF_EI,
    ORC_BUF_E_VAR, ORC_BUF_ETI, ORC_BUF_ET, /* SUBGRID_ECR */
    ORC_BUF_COUNT
};

typedef struct orc_sim orc_sim;

/* Returns NULL and fills err (if non-NULL, >= 256 bytes) on invalid params. */
orc_sim* orc_create(const orc_params* p, char* err);
void orc_destroy(orc_sim* s);
/* Raw host pointer, element count and element size of a buffer (NULL if the
 * buffer does not exist for this configuration). */
void* orc_buffer(orc_sim* s, int id, uint64_t* count, uint32_t* elem_size);
uint64_t orc_time_step(const orc_sim* s);
void orc_set_time_step(orc_sim* s, uint64_t t);
uint64_t orc_n_lod(const orc_sim* s);

/* Lbm::initialize / run / do_time_step (src/lbm/mod.rs:214-272). */
void orc_initialize(orc_sim* s);
void orc_run(orc_sim* s, uint64_t steps);
void orc_do_time_step(orc_sim* s);

/* Individual kernels at the current t, for unit tests. */
void orc_kernel_initialize(orc_sim* s);       /* sim_kernels.cl:760-832 */
void orc_kernel_stream_collide(orc_sim* s);   /* sim_kernels.cl:465-758 + LOD tree */
void orc_kernel_update_fields(orc_sim* s);    /* sim_kernels.cl:834-859 */
void orc_kernel_clear_qu_lod(orc_sim* s);     /* sim_kernels.cl:995-1003 */
void orc_kernel_update_e_b_dynamic(orc_sim* s); /* sim_kernels.cl:897-993 */

/* Float64 sums over fluid cells: mass sum(rho) and charge sum(Q). */
double orc_sum_rho(orc_sim* s);
double orc_sum_q(orc_sim* s);

/* Storage codecs (sim_kernels.cl:79-90, src/lbm/domain.rs:773-784). */
void orc_fp16c_encode(const float* x, uint16_t* h, uint64_t n);
void orc_fp16c_decode(const uint16_t* h, float* x, uint64_t n);
void orc_half_encode_rte(const float* x, uint16_t* h, uint64_t n); /* IEEE binary16, RNE */
void orc_half_decode(const uint16_t* h, float* x, uint64_t n);
void orc_fp16s_store(const float* x, uint16_t* h, uint64_t n); /* half_rte(x * 32768) */
void orc_fp16s_load(const uint16_t* h, float* x, uint64_t n);  /* half(h) * 2^-15 */

/* Per-cell lattice functions (velocity set vs), float arithmetic. */
void orc_f_eq(int vs, float rho, float ux, float uy, float uz, float* feq);
void orc_rho_u(int vs, const float* f, float* rho_u4);
void orc_forcing(int vs, float ux, float uy, float uz, float fx, float fy, float fz, float* fin);
void orc_a_eq(float q, float ux, float uy, float uz, float* qeq);
/* The same in double arithmetic (FP64 oracle). */
void orc_f_eq_f64(int vs, double rho, double ux, double uy, double uz, double* feq);
void orc_rho_u_f64(int vs, const double* f, double* rho_u4);

int orc_num_threads(void);
void orc_set_num_threads(int n);

/* Mesh voxelisation and static fields (mesh_oracle.cpp), on raw buffers:
 * voxelize_mesh (sim_kernels.cl:1150-1231) for the triangles tri[9 T] with
 * mesh bounds pmin/pmax (widened by 2 inside, src/mesh.rs:283-302); b_dyn
 * may be NULL (no MAGNETO_HYDRO). */
void orc_voxelize(uint32_t nx, uint32_t ny, uint32_t nz, uint8_t* flags, float* b_dyn, const float* tri, uint32_t T,
                  const float* pmin, const float* pmax, uint8_t flag, float mx, float my, float mz);
/* psi_from_mesh + static_b_from_mesh (:1234-1276): psi[(nx+2)(ny+2)(nz+2)],
 * B = B_stat, M = B_dyn. */
void orc_precompute_b(uint32_t nx, uint32_t ny, uint32_t nz, const uint8_t* flags, const float* M, float* psi,
                      float* B, float kmu0);
/* static_e_from_mesh (:1278-1299): E = E_stat, C = B_dyn. */
void orc_precompute_e(uint32_t nx, uint32_t ny, uint32_t nz, const uint8_t* flags, const float* C, float* E, float ke);

#ifdef __cplusplus
}
#endif
#endif
