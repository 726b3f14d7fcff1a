/*
 * ionsolver.h - C ABI of the MI355X (gfx950) backend of the IonSolver
 * extended-LBM MHD time step (libionsolver_hip.so).
 *
 * The ABI replaces the reference's Rust `Lbm`/`LbmConfig` API over its
 * OpenCL `LbmDomain` (src/lbm/mod.rs, src/lbm/domain.rs, src/opencl.rs)
 * for a single GPU. Every entry point cites the reference interface it
 * replaces. Conventions:
 *   - every function returns ION_OK (0) or a negative ion_status, and never
 *     aborts; ion_last_error() gives the message of the last failure on the
 *     calling thread;
 *   - one handle owns all of its device memory and one HIP stream; calls on
 *     one handle must be serialised by the caller;
 *   - host pointers are borrowed for the duration of a call; field copies are
 *     synchronous, like the reference's blocking bread!/bwrite!.
 */
#ifndef IONSOLVER_H
#define IONSOLVER_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ion_status {
    ION_OK = 0,
    ION_ERR_INVALID = -1,     /* invalid argument or configuration */
    ION_ERR_UNSUPPORTED = -2, /* valid for the reference, not for this backend */
    ION_ERR_DEVICE = -3,      /* HIP runtime error */
    ION_ERR_IO = -4,          /* file error */
    ION_ERR_FORMAT = -5,      /* malformed save file */
    ION_ERR_NONFINITE = -6    /* the watchdog found a non-finite mass or charge sum */
} ion_status;

/* VelocitySet, src/lbm/types.rs:16-26 (same discriminants) */
typedef enum { ION_D2Q9 = 0, ION_D3Q15 = 1, ION_D3Q19 = 2, ION_D3Q27 = 3 } ion_velocity_set;
/* RelaxationTime, src/lbm/types.rs:55-61 */
typedef enum { ION_SRT = 0, ION_TRT = 1 } ion_relaxation_time;
/* FloatType, src/lbm/types.rs:75-83 */
/* ION_FP64 has no reference counterpart: the FP64 instantiation of the same
 * kernels (double DDFs, double fields, double arithmetic, D3Q19 only), the
 * error reference of the storage formats. Its field buffers hold doubles
 * (ion_field_size reports 8-byte elements); it has no file format. */
typedef enum { ION_FP16S = 0, ION_FP16C = 1, ION_FP32 = 2, ION_FP64 = 3 } ion_float_type;
/* Propellant, src/lbm/units.rs:214-222 */
typedef enum { ION_PROP_H = 0, ION_PROP_HE, ION_PROP_NE, ION_PROP_AR, ION_PROP_KR, ION_PROP_XE } ion_propellant;

/* Units, src/lbm/units.rs:13-26: lattice unit * factor = SI unit. */
typedef struct ion_units {
    float m, kg, s, a, k;
    int32_t prop;
} ion_units;

/* LbmConfig, src/lbm/mod.rs:45-101. graphics_active stands for
 * graphics_config.graphics_active, the only GraphicsConfig field that reaches
 * the kernels (it switches UPDATE_FIELDS on, src/lbm/domain.rs:856).
 * ext_charge_inlet has no reference counterpart (the reference has no charge
 * injection): with ext_magneto_hydro and ext_equilibrium_boudaries it makes
 * the TYPE_E cells charge sources as well. There the charge density is taken
 * from the Q buffer (set by the caller, like rho and u of TYPE_E cells),
 * the D3Q7 charge DDFs are reset to their equilibrium with that charge every
 * step (no collision), and Q is not written back. It is not part of the save
 * format: ion_load takes it from cfg_inout, like the other fields the file
 * does not store. electron_density0 has no reference counterpart either:
 * initialize writes the electron-gas DDFs of non-solid cells as the
 * equilibrium of density electron_density0 (the reference writes density 0,
 * src/kernels/sim_kernels.cl:823-825, which leaves the electron density at the
 * rounding residue of the shifted sum); 0 (the default) is the reference. Not
 * in the save format either. */
typedef struct ion_config {
    int32_t velocity_set;
    int32_t relaxation_time;
    int32_t float_type;
    ion_units units;
    uint32_t n_x, n_y, n_z;
    uint32_t d_x, d_y, d_z;
    float nu;
    float f_x, f_y, f_z;
    uint8_t ext_equilibrium_boudaries;
    uint8_t ext_volume_force;
    uint8_t ext_force_field;
    uint8_t ext_magneto_hydro;
    uint8_t ext_subgrid_ecr;
    uint8_t mhd_lod_depth;
    uint8_t graphics_active;
    uint8_t ext_charge_inlet;
    float ecr_freq;
    float ecr_field_strength;
    uint64_t run_steps;
    float electron_density0;
} ion_config;

/* Device buffers of one domain (src/lbm/domain.rs:45-74), all SoA:
 * scalar fields [N] f32, vector fields [3][N] f32, flags [N] u8,
 * DDFs [q][N] (f32 for FP32, u16 codes for FP16S/FP16C), QU_lod [n_lod][4] f32. */
typedef enum ion_field {
    ION_FIELD_FI = 0,
    ION_FIELD_RHO = 1,
    ION_FIELD_U = 2,
    ION_FIELD_FLAGS = 3,
    ION_FIELD_F = 4,       /* force field (ext_force_field) */
    ION_FIELD_E_STAT = 5,
    ION_FIELD_B_STAT = 6,
    ION_FIELD_E_DYN = 7,
    ION_FIELD_B_DYN = 8,
    ION_FIELD_FQI = 9,     /* D3Q7 charge DDFs */
    ION_FIELD_Q = 10,
    ION_FIELD_QU_LOD = 11,
    ION_FIELD_EI = 12,     /* electron-gas DDFs */
    /* no reference counterpart: the raw source sums of the last direct field
     * update (E then B, 6 x N f32, before the KE / KMU scaling), kept once a
     * surrogate exists (the surrogate's training and judging reference) */
    ION_FIELD_FIELD_SUMS = 13,
    /* ext_subgrid_ecr (src/lbm/domain.rs:200-211) */
    ION_FIELD_E_VAR = 14,  /* oscillating electric field, 3 x N f32 */
    ION_FIELD_ETI = 15,    /* electron temperature DDFs (D3Q7), 7 x N storage codes */
    ION_FIELD_ET = 16,     /* electron temperature, N f32 */
    ION_FIELD_COUNT = 17
} ion_field;

/* The f32 kernel constants the reference bakes in as #defines
 * (src/lbm/domain.rs:808, 836-849), derived on the host in f32 in the
 * order of src/lbm/units.rs. */
typedef struct ion_kernel_constants {
    float w;     /* DEF_W  = 1/(3 nu + 1/2) */
    float c;     /* DEF_C  = 0.57735027 */
    float ke;    /* DEF_KE = 1/(4 pi eps0_lu) */
    float kmu;   /* DEF_KMU = mu0_lu/(4 pi) */
    float kmu0;  /* DEF_KMU0 = mu0_lu */
    float kkge;  /* DEF_KKGE = electron mass/charge in lattice units */
    float wq;    /* DEF_WQ = 1/(2 k + 1/2), k = 1 */
    uint32_t n_lod;     /* DEF_NUM_LOD (single domain: = n_lod_own) */
    float kkbme; /* DEF_KKBME = -1.5 k_B / m_e in lattice units (units.rs:187-189) */
    float keabs; /* DEF_KEABS (units.rs:191-193) */
} ion_kernel_constants;

/* Field-update arithmetic (build-side option, no reference counterpart):
 *   ION_FIELD_EXACT: the reference's r/|r|^3 with correctly rounded sqrt and
 *                    division and unfused accumulation (default);
 *   ION_FIELD_FAST:  hardware reciprocal square root and fused accumulation,
 *                    within a stated relative tolerance of EXACT. */
typedef enum { ION_FIELD_EXACT = 0, ION_FIELD_FAST = 1 } ion_field_mode;

/* Accumulated GPU time per kernel (hipEvent pairs on the handle's stream),
 * enabled by ion_set_timing. */
typedef struct ion_timing {
    double ms_stream_collide;
    double ms_lod_reduce;
    double ms_field_prep;
    double ms_update_e_b;
    double ms_other;
    uint64_t n_steps;
} ion_timing;

typedef struct ion_handle ion_handle;

/* LbmConfig::new, src/lbm/mod.rs:105-135 (graphics_active default true,
 * src/lbm/graphics.rs:160). */
void ion_config_default(ion_config* cfg);
/* Units::new / Units::set, src/lbm/units.rs:30-59. */
void ion_units_default(ion_units* u);
void ion_units_set(ion_units* u, float lbm_length, float lbm_velocity, float lbm_rho, float lbm_charge,
                   float lbm_temp, float si_length, float si_velocity, float si_rho, float si_charge, float si_temp);
/* get_device_defines constants, src/lbm/domain.rs:736-857, units.rs:149-177. */
int ion_kernel_constants_of(const ion_config* cfg, ion_kernel_constants* out);

/* Lbm::new, src/lbm/mod.rs:166-212 (single domain: d_x = d_y = d_z = 1).
 * Allocates and zero-fills the buffers as LbmDomain::new does
 * (src/lbm/domain.rs:151-211: rho = 1, everything else 0). */
int ion_create(const ion_config* cfg, ion_handle** out);
void ion_destroy(ion_handle* h);
int ion_get_config(const ion_handle* h, ion_config* out);

/* Lbm::initialize, src/lbm/mod.rs:214-231. */
int ion_initialize(ion_handle* h);
/* Lbm::run, src/lbm/mod.rs:235-246 (initialises lazily). */
int ion_run(ion_handle* h, uint64_t steps);
/* Lbm::do_time_step, src/lbm/mod.rs:250-272. */
int ion_do_time_step(ion_handle* h);
/* Lbm::get_time_step, src/lbm/mod.rs:492-494. */
uint64_t ion_get_time_step(const ion_handle* h);
/* LbmDomain::enqueue_update_fields, src/lbm/domain.rs:430-441. */
int ion_update_fields(ion_handle* h);
/* Blocks until the handle's stream is idle (Lbm::finish_queues, mod.rs:275-279). */
int ion_synchronize(ion_handle* h);

/* bwrite!/bread! on the domain buffers (e.g. src/setup.rs:302, src/file.rs:233).
 * `bytes` must equal the buffer size reported by ion_field_size. */
int ion_field_size(const ion_handle* h, int field, size_t* bytes, uint32_t* elem_size);
int ion_write_field(ion_handle* h, int field, const void* src, size_t bytes);
int ion_read_field(ion_handle* h, int field, void* dst, size_t bytes);
/* Volume force kernel arguments, src/lbm/domain.rs:422-425. */
int ion_set_volume_force(ion_handle* h, float fx, float fy, float fz);

/* file::write / file::read, src/file.rs:12-30 (FILE_LAYOUT.txt).
 * ion_load creates a handle from the file (not initialised, as the
 * reference's file::read); cfg_inout supplies the fields the file does not
 * store (graphics_active, units.k, units.prop, ...) and receives the result. */
int ion_save(ion_handle* h, const char* path);
int ion_load(const char* path, ion_config* cfg_inout, ion_handle** out);

/* Float64 conservation sums over fluid cells: sum(rho), sum(Q). */
int ion_conservation(ion_handle* h, double* mass, double* charge);
/* NaN/Inf watchdog (no reference counterpart; the reference has no failure
 * detection): with every > 0, ion_run and ion_do_time_step compute the
 * float64 mass and charge sums of ion_conservation after every step whose
 * counter t is a multiple of `every`, and stop with ION_ERR_NONFINITE at the
 * first non-finite one (the state of that step stays readable). 0 (the
 * default) switches it off. ion_watchdog_status: the step of the first
 * non-finite sum (UINT64_MAX if none), and the sums of the last check. */
int ion_set_watchdog(ion_handle* h, uint32_t every);
int ion_watchdog_status(const ion_handle* h, uint64_t* first_nonfinite_step, double* mass, double* charge);

int ion_set_field_mode(ion_handle* h, int mode);
int ion_set_timing(ion_handle* h, int enable);
int ion_get_timing(ion_handle* h, ion_timing* out);

/* Device storage codec on host arrays (diagnostics): dir 0 encodes
 * float -> u16 code, dir 1 decodes u16 -> float, for ION_FP16S/ION_FP16C. */
int ion_device_codec(int float_type, int dir, const void* src, void* dst, uint64_t n);

/* Device buffers of ion_update_e_b_noalloc, in the layouts ion_read_field
 * uses: flags, q [N]; u, e_stat, b_stat, e_dyn, b_dyn [3][N] (x block, y
 * block, z block); qu_lod [n_lod][4] (q, vx, vy, vz), as stream_collide
 * leaves it. q and u are read only when the near-field block holds more than
 * one cell (n_x / 2^(2^mhd_lod_depth) > 1); they may be NULL otherwise. */
typedef struct ion_field_buffers {
    const uint8_t* flags;
    const float* q;
    const float* u;
    const float* qu_lod;
    const float* e_stat;
    const float* b_stat;
    float* e_dyn;
    float* b_dyn;
} ion_field_buffers;

/* Most LOD sources ion_update_e_b_noalloc holds: 8^5, so mhd_lod_depth <= 5. */
#define ION_NOALLOC_MAX_SOURCES 32768u

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The field step of ion_compute_fields
 * (update_e_b_dynamic, src/kernels/sim_kernels.cl:897-993) on caller-owned
 * device buffers; E_dyn and B_dyn are bit-identical to ion_compute_fields on
 * the same inputs. It allocates nothing: the compacted LOD source list lives
 * in a static device array of ION_NOALLOC_MAX_SOURCES entries (1 MiB), one
 * per GPU, held until the process ends. cfg supplies the lattice, the LOD
 * depth (<= 5) and the units (KE, KMU); ext_magneto_hydro, FP32/FP16S/FP16C
 * (fields are float). field_mode: ION_FIELD_EXACT or ION_FIELD_FAST.
 * Enqueued on `stream` (a hipStream_t; NULL: the null stream) without
 * synchronising; on the current device. Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU race, so serialise them (one
 * stream, or events between streams). */
int ion_update_e_b_noalloc(const ion_config* cfg, const ion_field_buffers* buf, int field_mode, void* stream);

/* ------------------------------------------------------------------------
 * Learned field surrogate (build-side; the reference has no counterpart).
 * A linear 4-level 3D conv neural operator on the LOD grid (BF16 MFMA, FP32
 * accumulation) that maps the top-level LOD charge and current sums to
 * E - E_stat and B - B_stat at every cell. The direct update_e_b_dynamic
 * is its judge: with ION_SOLVER_SURROGATE the direct kernel runs every
 * `judge_every` field steps, the relative L2 residual of the surrogate is
 * measured against it, and a residual above `tolerance` switches the next
 * `judge_every` field steps back to the direct kernel (a fallback).
 * Needs ext_magneto_hydro, mhd_lod_depth >= 3 and an empty near-field block.
 * ------------------------------------------------------------------------ */
typedef enum { ION_SOLVER_DIRECT = 0, ION_SOLVER_SURROGATE = 1 } ion_field_solver;

typedef struct ion_surrogate_config {
    int32_t channels;  /* level-0 channels c0; levels use c0, 2c0, 3c0, 4c0 (default 16) */
    uint64_t seed;     /* weight initialisation */
    float lr;          /* Adam step of the U-Net layers and the head's feature columns (default 1e-3) */
    float lr_head;     /* Adam step of the head's direct-source (near field) columns (default 1e-4) */
} ion_surrogate_config;

typedef struct ion_surrogate_stats {
    uint64_t train_steps;
    uint64_t direct_steps;     /* field steps solved by the direct kernel */
    uint64_t surrogate_steps;  /* field steps solved by the surrogate */
    uint64_t fallbacks;        /* judge residuals above the tolerance */
    double last_residual_e;    /* relative L2 error of the dynamic part E - E_stat at the last judge step */
    double last_residual_b;
} ion_surrogate_stats;

void ion_surrogate_config_default(ion_surrogate_config* cfg);
int ion_surrogate_create(ion_handle* h, const ion_surrogate_config* cfg);
/* LbmDomain::enqueue_update_e_b_dyn (src/lbm/domain.rs:443): direct field
 * update from the current LOD sums, Q and u. Once a surrogate exists it also
 * keeps the raw source sums (the dynamic parts before the KE / KMU scaling),
 * which are the surrogate's training and judging reference; E_dyn / B_dyn
 * are bit-identical either way. */
int ion_compute_fields(ion_handle* h);
/* Surrogate field from the current LOD sums into E_dyn / B_dyn. */
int ion_surrogate_infer(ion_handle* h);
/* One training step against the raw sums of the last ion_compute_fields
 * (call it first); update = 0 computes gradients only. rel_e / rel_b:
 * relative L2 errors of the dynamic parts before the step. */
int ion_surrogate_train_step(ion_handle* h, int update, double* rel_e, double* rel_b);
/* Relative L2 errors of the surrogate's dynamic E and B against the raw sums
 * of the last ion_compute_fields. */
int ion_surrogate_residual(ion_handle* h, double* rel_e, double* rel_b);
/* Adam steps for the following train steps (schedules). */
int ion_surrogate_set_learning_rate(ion_handle* h, float lr, float lr_head);
int ion_set_field_solver(ion_handle* h, int solver, uint32_t judge_every, double tolerance);
int ion_surrogate_get_stats(ion_handle* h, ion_surrogate_stats* out);
/* Flat fp32 parameters (or gradients of the last train step): layers
 * e0, e1, e2, e3, d2, d1, d0, head, each [cout][kpad] row-major. */
int ion_surrogate_param_count(ion_handle* h, uint64_t* n);
int ion_surrogate_read_params(ion_handle* h, int grads, float* dst, uint64_t n);
int ion_surrogate_write_params(ion_handle* h, const float* src, uint64_t n);

/* ------------------------------------------------------------------------
 * Measurement helper (no reference counterpart): copies `bytes` between two
 * fresh device buffers with a 16-byte nontemporal streaming kernel, reps + 1
 * times, and reports the best time and the read + write rate in GB/s. The
 * benchmark prints it beside the published HBM peak, and its known traffic
 * scales the FETCH_SIZE / WRITE_SIZE counters of profiles.
 * ------------------------------------------------------------------------ */
int ion_hbm_copy_probe(uint64_t bytes, uint32_t reps, double* best_ms, double* gb_per_s);
/* Measurement helper (no reference counterpart): stream_collide's access
 * shape without its arithmetic. One work-item per cell reads `streams`
 * (19 or 45) structure-of-arrays elements of `elem_bytes` (2 or 4), then
 * writes them to a second buffer at the same positions, nontemporal; best of
 * reps + 1 runs, rate in GB/s (read + write). */
int ion_hbm_stream_probe(uint32_t elem_bytes, uint32_t streams, uint32_t cells, uint32_t reps, double* best_ms,
                         double* gb_per_s);
/* Measurement helper (no reference counterpart): stream_collide's in-place
 * esoteric-pull addressing without its arithmetic, on an nx x ny x nz lattice.
 * `streams` (19: fi; 45: fi, ei, fqi) SoA arrays of `elem_bytes` (2 or 4);
 * one work-item per cell reads every array at the cell or, for the odd
 * directions i, at the periodic neighbour n + c_i (the layout of an even
 * step), then rewrites the same slots; best of reps + 1 runs, rate in GB/s
 * (read + write). */
int ion_hbm_pull_probe(uint32_t elem_bytes, uint32_t streams, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t reps,
                       double* best_ms, double* gb_per_s);

/* ------------------------------------------------------------------------
 * Meshes and static fields (src/mesh.rs, src/lbm/mod.rs:284-314).
 * Meshes belong to the handle; `index` is the position in Lbm::meshes.
 * ------------------------------------------------------------------------ */
/* ModelType, src/mesh.rs:9-14; flags 0x01, 0x11, 0x09, 0x05 (mesh.rs:309-314). */
typedef enum {
    ION_MODEL_SOLID = 0,
    ION_MODEL_MAGNET = 1,      /* magnetisation (vx, vy, vz) in A/m */
    ION_MODEL_CHARGED = 2,     /* charge vx in C per cell */
    ION_MODEL_CHARGED_ECR = 3  /* charge vx in C per cell */
} ion_model_type;

/* Lbm::import_mesh, src/mesh.rs:234-240: binary STL read, rotated by
 * (rx, ry, rz) degrees, scaled by len_si_lu(scale), placed at origin
 * (ox, oy, oz). ION_ERR_IO if unreadable, ION_ERR_FORMAT if not a binary
 * STL of 84 + 50 T bytes with T > 0 (the reference panics). */
int ion_import_mesh(ion_handle* h, const char* path, float scale, float ox, float oy, float oz, float rx, float ry,
                    float rz, uint32_t* index);
/* Lbm::import_mesh_reposition, src/mesh.rs:248-254: centred at (cx, cy, cz),
 * longest side `size` cells (0: fit the lattice box, < 0: scale factor -size). */
int ion_import_mesh_reposition(ion_handle* h, const char* path, float cx, float cy, float cz, float rx, float ry,
                               float rz, float size, uint32_t* index);
/* The imported triangles (vertices[9 T]: p0, p1, p2) and bounds
 * (p_min, p_max, center: 9 floats); either pointer may be null. */
int ion_mesh_get(ion_handle* h, uint32_t index, uint32_t* triangles, float* vertices, float* bounds);
/* Lbm::voxelise_mesh, src/mesh.rs:256-343 and voxelize_mesh,
 * src/kernels/sim_kernels.cl:1150-1231: cells inside the mesh get
 * flags = (flags & 0xE0) | flag of the model type; with ext_magneto_hydro a
 * magnet writes its magnetisation (lattice units) into B_dyn and a charged
 * model its charge into B_dyn.x, where precompute reads them. Run before
 * ion_initialize (which sets B_dyn = B_stat). */
int ion_voxelise_mesh(ion_handle* h, uint32_t index, int model_type, float vx, float vy, float vz);
/* Lbm::precompute_B, src/lbm/mod.rs:284-297 (psi_from_mesh +
 * static_b_from_mesh): B_stat += -mu0 grad psi of the magnetised cells.
 * psi is left in E_dyn (padded lattice) as in the reference. Synchronous. */
int ion_precompute_b(ion_handle* h);
/* Lbm::precompute_E, src/lbm/mod.rs:302-314 (static_e_from_mesh): E_stat +=
 * Coulomb field of the charged cells. Synchronous. */
int ion_precompute_e(ion_handle* h);
/* Lbm::precompute_E_ECR, src/lbm/mod.rs:319-331: the same field into E_var
 * (ext_subgrid_ecr). Synchronous. */
int ion_precompute_e_ecr(ion_handle* h);
/* Diagnostics (no reference counterpart): the specialised correctly rounded
 * square root and division of the precompute_B kernel against the general
 * ones on the current device: every integer in [1, 2^24) for the square root,
 * div_samples pseudo-random operand pairs of its range for the division. */
int ion_psi_math_check(uint64_t* sqrt_mismatch, uint64_t* div_mismatch, uint64_t* div_samples);

/* Message of the last failure on this thread ("" if none). */
const char* ion_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
