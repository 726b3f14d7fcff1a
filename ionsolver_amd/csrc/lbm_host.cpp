// Host runtime of the gfx950 backend: the C ABI of include/ionsolver.h.
//
// Mirrors the single-domain part of the reference host: LbmConfig and its
// defaults (src/lbm/mod.rs:45-135), Units (src/lbm/units.rs), buffer
// allocation (src/lbm/domain.rs:151-211), the kernel constants of
// get_device_defines (src/lbm/domain.rs:736-857), the step sequencing of
// Lbm::initialize / run / do_time_step (src/lbm/mod.rs:214-272) and the save
// format of src/file.rs / FILE_LAYOUT.txt. One handle = one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"
#include "ionsolver.h"
#include "kernels.hpp"
#include "surrogate_host.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                        \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return fail(ION_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// Units, src/lbm/units.rs (all f32 arithmetic in the reference's order)
// ---------------------------------------------------------------------------
inline float sq(float x) { return x * x; }
inline float cb(float x) { return x * x * x; }
inline float to4(float x) { return x * x * x * x; }
constexpr float PI_F = 3.14159265358979323846f;  // std::f32::consts::PI

float epsilon_0_lu(const ion_units& u) {  // units.rs:149-153
    return 8.8541878128E-12f / ((sq(u.a) * to4(u.s)) / (u.kg * cb(u.m)));
}
float ke_lu(const ion_units& u) { return 1.0f / (4.0f * PI_F * epsilon_0_lu(u)); }        // units.rs:155-161
float mu_0_lu(const ion_units& u) { return 1.256637062E-6f / ((u.kg * u.m) / (sq(u.a) * sq(u.s))); }  // :163-167
float kkge_lu(const ion_units& u) {  // units.rs:175-177: f64 quotient cast to f32, then f32
    return (float)(9.1093837139E-31 / -1.602176634E-19) / (u.kg / (u.a * u.s));
}

uint64_t n_lod_own_of(uint32_t D) {  // src/lbm/domain.rs:110-117 (3D, one domain)
    uint64_t c = 1;
    for (uint32_t i = 0; i < D; i++) c += (uint64_t)1 << (3 * (i + 1));
    return c;
}

int q_of(int vs) { return vs == ION_D2Q9 ? 9 : (vs == ION_D3Q15 ? 15 : (vs == ION_D3Q19 ? 19 : 27)); }

}  // namespace

struct ion_handle {
    ion_config cfg{};
    ion_kernel_constants kc{};
    int Q = 19;
    int store = ion::STORE_FP32;
    bool mhd = false, lod = false;
    uint32_t D = 0, nd = 1, lod_bx = 1;
    uint64_t N = 0, n_lod = 0;
    uint32_t sc_block = 256;
    hipStream_t stream = nullptr;
    // device buffers (src/lbm/domain.rs:45-74)
    void* fi = nullptr;
    void* ei = nullptr;
    void* fqi = nullptr;
    float *rho = nullptr, *u = nullptr, *F = nullptr, *e_stat = nullptr, *b_stat = nullptr, *e_dyn = nullptr,
          *b_dyn = nullptr, *q = nullptr, *qu_lod = nullptr;
    uint8_t* flags = nullptr;
    float4* lod_seg = nullptr;
    float4* src_pq = nullptr;
    float4* src_vd = nullptr;
    uint32_t* src_count = nullptr;
    double* partials = nullptr;
    uint32_t n_partials = 1024;
    std::vector<void*> allocs;
    uint64_t t = 0;
    bool initialized = false;
    float fx = 0, fy = 0, fz = 0;
    int field_mode = ION_FIELD_EXACT;
    // NaN/Inf watchdog
    uint32_t watch_every = 0;
    uint64_t watch_first_bad = UINT64_MAX;
    double watch_mass = 0.0, watch_charge = 0.0;
    // timing
    bool timing = false;
    ion_timing tm{};
    std::vector<hipEvent_t> ev_pool;
    struct Pend { hipEvent_t a, b; int slot; };
    std::vector<Pend> pending;
    // learned field surrogate and its judge
    ion::sur::Net* net = nullptr;
    float* field_raw = nullptr;  // raw direct source sums (E then B, 6 x N) once a surrogate exists
    double* judge_partials = nullptr;  // the judge's residual sums [4] then [judge_blocks][4]
    uint32_t judge_blocks = 0;
    int solver = ION_SOLVER_DIRECT;
    uint32_t judge_every = 50;
    double tolerance = 0.01;
    uint64_t field_steps = 0;
    uint32_t fallback_left = 0;
    ion_surrogate_stats sstats{};
    // imported meshes (mesh_host.cpp)
    ion::MeshSet* meshes = nullptr;
    // SUBGRID_ECR (domain.rs:200-211)
    bool ecr = false;
    void* eti = nullptr;
    float *et = nullptr, *e_var = nullptr;

    size_t ddf_elem() const { return store == ion::STORE_FP64 ? 8 : (store == ion::STORE_FP32 ? 4 : 2); }
    bool f64() const { return store == ion::STORE_FP64; }
    size_t real() const { return f64() ? 8 : 4; }  // element size of the non-DDF float fields
};

namespace {

int dalloc(ion_handle* h, void** p, size_t bytes) {
    if (bytes == 0) { *p = nullptr; return ION_OK; }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(ION_ERR_DEVICE, "hipMalloc(" + std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
    h->allocs.push_back(*p);
    e = hipMemsetAsync(*p, 0, bytes, h->stream);
    if (e != hipSuccess) return fail(ION_ERR_DEVICE, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    return ION_OK;
}

void free_all(ion_handle* h) {
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h->net;
    h->net = nullptr;
    ion::mesh_set_destroy(h->meshes);
    h->meshes = nullptr;
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    for (auto& p : h->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    h->ev_pool.clear();
    h->pending.clear();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    h->stream = nullptr;
}

hipEvent_t take_event(ion_handle* h) {
    if (!h->ev_pool.empty()) {
        hipEvent_t e = h->ev_pool.back();
        h->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

int drain_timing(ion_handle* h) {
    if (h->pending.empty()) return ION_OK;
    HIPCHK(hipStreamSynchronize(h->stream));
    for (auto& p : h->pending) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
        double* slot[] = {&h->tm.ms_stream_collide, &h->tm.ms_lod_reduce, &h->tm.ms_field_prep, &h->tm.ms_update_e_b,
                          &h->tm.ms_other};
        *slot[p.slot] += ms;
        h->ev_pool.push_back(p.a);
        h->ev_pool.push_back(p.b);
    }
    h->pending.clear();
    return ION_OK;
}

// Runs `launch` on the handle's stream, bracketed by events when timing.
template <class L>
int timed(ion_handle* h, int slot, L&& launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (h->timing) {
        a = take_event(h);
        HIPCHK(hipEventRecord(a, h->stream));
    }
    hipError_t e = launch();
    if (e != hipSuccess) return fail(ION_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    if (h->timing) {
        b = take_event(h);
        HIPCHK(hipEventRecord(b, h->stream));
        h->pending.push_back({a, b, slot});
        if (h->pending.size() >= 4096) return drain_timing(h);
    }
    return ION_OK;
}

ion::LbmArgs lbm_args(ion_handle* h) {
    ion::LbmArgs a{};
    a.fi = h->fi;
    a.rho = h->rho;
    a.u = h->u;
    a.flags = h->flags;
    a.F = h->F;
    a.e_stat = h->e_stat;
    a.b_stat = h->b_stat;
    a.e_dyn = h->e_dyn;
    a.b_dyn = h->b_dyn;
    a.fqi = h->fqi;
    a.ei = h->ei;
    a.q = h->q;
    a.lod_seg = h->lod_seg;
    a.nx = h->cfg.n_x;
    a.ny = h->cfg.n_y;
    a.nz = h->cfg.n_z;
    a.N = h->N;
    a.N_store = h->N;
    a.odd = (uint32_t)(h->t & 1u);
    a.eqb = h->cfg.ext_equilibrium_boudaries ? 1u : 0u;
    a.vforce = h->cfg.ext_volume_force ? 1u : 0u;
    a.ffield = h->cfg.ext_force_field ? 1u : 0u;
    a.upd = h->cfg.graphics_active ? 1u : 0u;
    a.qin = h->cfg.ext_charge_inlet ? 1u : 0u;
    a.ei_rho0 = h->cfg.electron_density0;
    a.eti = h->eti;
    a.et = h->et;
    a.e_var = h->e_var;
    a.kkbme = h->kc.kkbme;
    a.keabs = h->kc.keabs;
    a.lod_bx = h->lod_bx;
    a.fx = h->fx;
    a.fy = h->fy;
    a.fz = h->fz;
    a.w = h->kc.w;
    a.c = h->kc.c;
    a.kkge = h->kc.kkge;
    a.wq = h->kc.wq;
    // ils = 1.0f/lod_s(D), lod_s = (float)(bx*by*bz), sim_kernels.cl:436-440, 672
    const uint32_t nd = h->nd;
    a.ils = h->lod ? 1.0f / (float)((uint64_t)(h->cfg.n_x / nd) * (h->cfg.n_y / nd) * (h->cfg.n_z / nd)) : 0.0f;
    a.ils_d = h->lod ? 1.0 / (double)(float)((uint64_t)(h->cfg.n_x / nd) * (h->cfg.n_y / nd) * (h->cfg.n_z / nd)) : 0.0;
    return a;
}

// update_e_b_dynamic for the current buffers (sim_kernels.cl:897-993).
// Compacted LOD sources (q != 0) of the range update_e_b_dynamic reads.
int prep_sources(ion_handle* h) {
    const uint64_t n_top = (uint64_t)h->nd * h->nd * h->nd;
    const uint32_t hi = (uint32_t)h->n_lod, lo = (uint32_t)(h->n_lod - n_top);
    return timed(h, 2, [&] {
        return ion::launch_prep_sources(h->qu_lod, h->f64(), lo, hi, h->cfg.n_x, h->cfg.n_y, h->cfg.n_z, h->D,
                                        h->src_pq, h->src_vd, h->src_count, h->stream);
    });
}

ion::FieldArgs field_args(ion_handle* h) {
    ion::FieldArgs f{};
    f.e_stat = h->e_stat;
    f.b_stat = h->b_stat;
    f.e_dyn = h->e_dyn;
    f.b_dyn = h->b_dyn;
    f.flags = h->flags;
    f.q = h->q;
    f.u = h->u;
    f.src_pq = h->src_pq;
    f.src_vd = h->src_vd;
    f.src_count = h->src_count;
    f.nx = h->cfg.n_x;
    f.ny = h->cfg.n_y;
    f.nz = h->cfg.n_z;
    f.D = h->D;
    f.N = h->N;
    // near-field block side imax(N/(1<<nd), 1); OpenCL masks the shift count
    // of a 32-bit int to 5 bits (sim_kernels.cl:906-910)
    const uint32_t sh = 1u << (h->nd & 31u);
    auto side = [&](uint32_t n) { uint32_t v = n / sh; return v < 1 ? 1u : v; };
    f.dsx = side(h->cfg.n_x);
    f.dsy = side(h->cfg.n_y);
    f.dsz = side(h->cfg.n_z);
    f.near = ((uint64_t)f.dsx * f.dsy * f.dsz > 1) ? 1u : 0u;
    f.ke = h->kc.ke;
    f.kmu = h->kc.kmu;
    f.raw = h->field_raw;
    f.f64 = h->f64() ? 1u : 0u;
    return f;
}

int field_update(ion_handle* h) {
    int rc = prep_sources(h);
    if (rc) return rc;
    const ion::FieldArgs f = field_args(h);
    return timed(h, 3, [&] {
        hipError_t e = ion::launch_update_e_b(f, h->field_mode == ION_FIELD_FAST, h->stream);
        if (e == hipSuccess && f.raw != nullptr) e = ion::launch_combine_fields(f, h->stream);
        return e;
    });
}


ion::sur::Fields sur_fields(ion_handle* h) {
    return ion::sur::Fields{(const float4*)h->qu_lod, h->flags, h->e_stat, h->b_stat, h->e_dyn, h->b_dyn,
                            h->field_raw, h->src_pq, h->src_vd, h->src_count};
}

// The judge's field step: the surrogate writes its raw source sums into
// E_dyn / B_dyn, then the direct kernel compares its own sums with them while
// it overwrites them with the direct fields (bit-identical to a direct step),
// so no 6 x N buffer is needed for either. Residuals: relative L2 of the
// dynamic E and B over the fluid cells.
int judged_field_update(ion_handle* h, double* rel_e, double* rel_b) {
    int rc = prep_sources(h);
    if (rc) return rc;
    ion::FieldArgs f = field_args(h);
    f.raw = nullptr;
    const bool fast = h->field_mode == ION_FIELD_FAST;
    if (!h->judge_partials) {
        const uint32_t nb = std::max(ion::field_blocks(f, true), ion::field_blocks(f, false));
        if ((rc = dalloc(h, (void**)&h->judge_partials, ((size_t)nb + 1) * 4 * sizeof(double)))) return rc;
        h->judge_blocks = nb;
    }
    f.judge = h->judge_partials + 4;
    bool ok = true;
    rc = timed(h, 3, [&] {
        ok = h->net->infer_raw(sur_fields(h)) == 0;
        if (!ok) return hipSuccess;
        hipError_t e = ion::launch_update_e_b(f, fast, h->stream);
        if (e == hipSuccess) e = ion::launch_judge_reduce(f.judge, ion::field_blocks(f, fast), h->judge_partials, h->stream);
        return e;
    });
    if (rc) return rc;
    if (!ok) return fail(ION_ERR_DEVICE, "surrogate: " + h->net->error());
    double s[4];
    HIPCHK(hipMemcpyAsync(s, h->judge_partials, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *rel_e = s[1] > 0 ? std::sqrt(s[0] / s[1]) : 0.0;
    *rel_b = s[3] > 0 ? std::sqrt(s[2] / s[3]) : 0.0;
    return ION_OK;
}

// The field step of a time step: the direct kernel, or the surrogate with the
// direct kernel as its judge every judge_every field steps (and for
// judge_every steps after a residual above the tolerance).
int field_step(ion_handle* h) {
    if (h->solver != ION_SOLVER_SURROGATE || !h->net) {
        h->sstats.direct_steps += 1;
        return field_update(h);
    }
    const bool judge = h->fallback_left > 0 || h->field_steps % h->judge_every == 0;
    h->field_steps += 1;
    int rc;
    if (judge) {
        double re = 0, rb = 0;
        if ((rc = judged_field_update(h, &re, &rb))) return rc;
        h->sstats.direct_steps += 1;
        if (h->fallback_left > 0) h->fallback_left -= 1;
        h->sstats.last_residual_e = re;
        h->sstats.last_residual_b = rb;
        if (!(re <= h->tolerance && rb <= h->tolerance)) {  // a NaN residual is a miss too
            h->sstats.fallbacks += 1;
            h->fallback_left = h->judge_every;
        }
        return ION_OK;
    }
    if ((rc = prep_sources(h))) return rc;
    bool ok = true;
    rc = timed(h, 3, [&] {
        ok = h->net->infer(sur_fields(h)) == 0;
        return hipSuccess;
    });
    if (rc) return rc;
    if (!ok) return fail(ION_ERR_DEVICE, "surrogate: " + h->net->error());
    h->sstats.surrogate_steps += 1;
    return ION_OK;
}

int do_step(ion_handle* h) {
    int rc;
    if (h->mhd) {  // clear_qu_lod, sim_kernels.cl:995-1003
        rc = timed(h, 4, [&] { return ion::launch_fill(h->qu_lod, h->f64(), h->n_lod * 4, 0.0, h->stream); });
        if (rc) return rc;
    }
    ion::LbmArgs a = lbm_args(h);
    rc = timed(h, 0, [&] {
        return ion::launch_stream_collide(h->Q, h->store, h->mhd, h->cfg.relaxation_time == ION_TRT, h->lod,
                                          h->sc_block, a, h->stream, h->ecr);
    });
    if (rc) return rc;
    if (h->mhd) {
        if (h->lod) {
            rc = timed(h, 1, [&] {
                return ion::launch_lod_reduce(h->lod_seg, h->qu_lod, h->f64(), h->cfg.n_x, h->cfg.n_y, h->cfg.n_z, h->D,
                                              h->stream);
            });
            if (rc) return rc;
        }
        rc = field_step(h);
        if (rc) return rc;
    }
    h->t += 1;
    if (h->timing) h->tm.n_steps += 1;
    return ION_OK;
}

int validate(const ion_config& c) {
    if (c.velocity_set == ION_D3Q27)
        return fail(ION_ERR_UNSUPPORTED, "D3Q27 is not supported: the reference defines no DEF_WC for it (domain.rs:754-756)");
    if (c.velocity_set < ION_D2Q9 || c.velocity_set > ION_D3Q27) return fail(ION_ERR_INVALID, "invalid velocity_set");
    if (c.relaxation_time != ION_SRT && c.relaxation_time != ION_TRT) return fail(ION_ERR_INVALID, "invalid relaxation_time");
    if (c.float_type < ION_FP16S || c.float_type > ION_FP64) return fail(ION_ERR_INVALID, "invalid float_type");
    if (c.float_type == ION_FP64 && c.velocity_set != ION_D3Q19)
        return fail(ION_ERR_UNSUPPORTED, "the FP64 instantiation is built for D3Q19 only");
    if (c.n_x == 0 || c.n_y == 0 || c.n_z == 0) return fail(ION_ERR_INVALID, "lattice size must be positive");
    if (c.d_x != 1 || c.d_y != 1 || c.d_z != 1)
        return fail(ION_ERR_UNSUPPORTED, "multi-domain split (d_x, d_y, d_z != 1) is replaced by a single GPU");
    // n_x * n_y first, so that the product of three 32-bit sides cannot wrap
    const uint64_t nxy = (uint64_t)c.n_x * c.n_y;
    if (nxy > 0x40000000ull || nxy * c.n_z > 0x40000000ull)
        return fail(ION_ERR_UNSUPPORTED, "more than 2^30 cells (1024^3) on one GPU: the kernels address every "
                                         "per-direction array with 32-bit byte offsets");
    const uint64_t N = nxy * c.n_z;
    if (c.float_type == ION_FP64 && N > 0x20000000ull)
        return fail(ION_ERR_UNSUPPORTED, "more than 2^29 cells with FP64: the kernels address every per-direction "
                                         "array with 32-bit byte offsets, and 8-byte elements reach 2^32 bytes there");
    if (c.ext_subgrid_ecr) {
        if (!c.ext_magneto_hydro) return fail(ION_ERR_INVALID, "ext_subgrid_ecr needs ext_magneto_hydro");
        if (c.velocity_set != ION_D3Q19 || c.float_type == ION_FP64)
            return fail(ION_ERR_UNSUPPORTED, "ext_subgrid_ecr is built for D3Q19 with FP32, FP16S or FP16C storage");
    }
    if (!(c.electron_density0 >= 0.0f) || !std::isfinite(c.electron_density0))
        return fail(ION_ERR_INVALID, "electron_density0 must be finite and >= 0");
    if (c.electron_density0 != 0.0f && !c.ext_magneto_hydro)
        return fail(ION_ERR_INVALID, "electron_density0 needs ext_magneto_hydro");
    if (c.ext_force_field && !c.ext_volume_force) return fail(ION_ERR_INVALID, "ext_force_field needs ext_volume_force");
    if (c.ext_charge_inlet && !(c.ext_magneto_hydro && c.ext_equilibrium_boudaries))
        return fail(ION_ERR_INVALID, "ext_charge_inlet needs ext_magneto_hydro and ext_equilibrium_boudaries");
    if (c.ext_magneto_hydro) {
        if (c.velocity_set == ION_D2Q9) return fail(ION_ERR_UNSUPPORTED, "ext_magneto_hydro needs a 3D velocity set");
        if (!c.ext_volume_force) return fail(ION_ERR_INVALID, "ext_magneto_hydro needs ext_volume_force");
        if (c.mhd_lod_depth > 7) return fail(ION_ERR_UNSUPPORTED, "mhd_lod_depth > 7");
        if (c.mhd_lod_depth > 0) {
            const uint32_t nd = 1u << c.mhd_lod_depth;
            if (c.n_x % nd || c.n_y % nd || c.n_z % nd)
                return fail(ION_ERR_UNSUPPORTED, "ext_magneto_hydro needs n_x, n_y, n_z divisible by 2^mhd_lod_depth");
        }
    }
    return ION_OK;
}

// stream_collide block: a contiguous range of whole LOD row segments when
// the level-1 LOD tree is fused in (whole rows up to 256 cells, else 256
// cells of a row where the segments allow, at most 512), else 256 cells.
int choose_block(ion_handle* h) {
    if (!h->lod) { h->sc_block = 256; return ION_OK; }
    const uint32_t NX = h->cfg.n_x, bx = h->lod_bx;
    const uint64_t rows = (uint64_t)h->cfg.n_y * h->cfg.n_z;
    if (NX <= 256) {
        uint32_t ty = 1;
        for (uint32_t t = 2; (uint64_t)NX * t <= 256; t++)
            if (rows % t == 0) ty = t;
        h->sc_block = NX * ty;
        return ION_OK;
    }
    // wider rows: part of a row, whole LOD segments, preferably 256 cells
    for (uint32_t cap : {256u, 512u})
        for (uint32_t b = cap; b >= bx && b > 0; b--)
            if (NX % b == 0 && b % bx == 0) { h->sc_block = b; return ION_OK; }
    return fail(ION_ERR_UNSUPPORTED, "LOD block wider than 512 cells along x");
}

}  // namespace

extern "C" {

const char* ion_last_error(void) { return g_err.c_str(); }

__attribute__((visibility("hidden"))) int ion_internal_set_error(int code, const char* msg) { return fail(code, msg); }

__attribute__((visibility("hidden"))) int ion_internal_view(ion_handle* h, ion::HandleView* v) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    v->stream = h->stream;
    v->nx = h->cfg.n_x;
    v->ny = h->cfg.n_y;
    v->nz = h->cfg.n_z;
    v->N = h->N;
    v->mhd = h->mhd;
    v->f64 = h->f64();
    v->cfg = &h->cfg;
    v->kc = &h->kc;
    v->flags = h->flags;
    v->e_stat = h->e_stat;
    v->b_stat = h->b_stat;
    v->e_dyn = h->e_dyn;
    v->b_dyn = h->b_dyn;
    v->meshes = &h->meshes;
    v->e_var = h->e_var;
    return ION_OK;
}

void ion_units_default(ion_units* u) {
    u->m = 1.0f; u->kg = 1.0f; u->s = 1.0f; u->a = 1.0f; u->k = 1.0f;
    u->prop = ION_PROP_H;
}

void ion_units_set(ion_units* u, float lbm_length, float lbm_velocity, float lbm_rho, float lbm_charge, float lbm_temp,
                   float si_length, float si_velocity, float si_rho, float si_charge, float si_temp) {
    // Units::set, src/lbm/units.rs:41-59
    u->m = si_length / lbm_length;
    u->kg = si_rho / lbm_rho * cb(u->m);
    u->s = u->m / (si_velocity / lbm_velocity);
    u->a = si_charge / lbm_charge / u->s;
    u->k = si_temp / lbm_temp;
}

void ion_config_default(ion_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->velocity_set = ION_D2Q9;
    c->relaxation_time = ION_SRT;
    c->float_type = ION_FP16S;
    ion_units_default(&c->units);
    c->n_x = c->n_y = c->n_z = 1;
    c->d_x = c->d_y = c->d_z = 1;
    c->nu = 1.0f / 6.0f;
    c->mhd_lod_depth = 4;
    c->graphics_active = 1;
}

int ion_kernel_constants_of(const ion_config* c, ion_kernel_constants* k) {
    if (!c || !k) return fail(ION_ERR_INVALID, "null argument");
    k->w = 1.0f / (3.0f * c->nu + 0.5f);  // domain.rs:808
    k->c = 0.57735027f;
    k->ke = ke_lu(c->units);
    k->kmu0 = mu_0_lu(c->units);
    k->kmu = k->kmu0 / (4.0f * PI_F);
    k->kkge = kkge_lu(c->units);
    const float kce = 1.0f;  // k_charge_expansion_lu, units.rs:169-173
    k->wq = 1.0f / (2.0f * kce + 0.5f);
    k->n_lod = c->ext_magneto_hydro ? (uint32_t)n_lod_own_of(c->mhd_lod_depth) : 0u;
    // units.rs:187-193: f64 quotients cast to f32
    const double m = c->units.m, s = c->units.s, kk = c->units.k, a = c->units.a, kg = c->units.kg;
    k->kkbme = (float)(-22734499.72063751808909449412 / ((m * m) / ((s * s) * kk)));
    k->keabs = (float)(1.40897016100511360652E-8 / ((a * a) * (s * s) / kg));
    return ION_OK;
}

int ion_create(const ion_config* cfg, ion_handle** out) {
    if (!cfg || !out) return fail(ION_ERR_INVALID, "null argument");
    *out = nullptr;
    int rc = validate(*cfg);
    if (rc) return rc;
    ion_handle* h = new ion_handle();
    h->cfg = *cfg;
    ion_kernel_constants_of(cfg, &h->kc);
    h->Q = q_of(cfg->velocity_set);
    h->store = cfg->float_type;
    h->N = (uint64_t)cfg->n_x * cfg->n_y * cfg->n_z;
    h->mhd = cfg->ext_magneto_hydro != 0;
    h->fx = cfg->f_x;
    h->fy = cfg->f_y;
    h->fz = cfg->f_z;
    if (h->mhd) {
        h->D = cfg->mhd_lod_depth;
        h->nd = 1u << h->D;
        h->n_lod = n_lod_own_of(h->D);
        h->lod = h->D > 0;
        h->lod_bx = h->lod ? cfg->n_x / h->nd : 1;
    }
    auto bail = [&](int code) {
        free_all(h);
        delete h;
        return code;
    };
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { fail(ION_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e)); return bail(ION_ERR_DEVICE); }
    if ((rc = choose_block(h))) return bail(rc);
    const uint64_t N = h->N;
    const size_t de = h->ddf_elem();
    if ((rc = dalloc(h, &h->fi, N * h->Q * de))) return bail(rc);
    const size_t rs = h->real();
    if ((rc = dalloc(h, (void**)&h->rho, N * rs))) return bail(rc);
    if ((rc = dalloc(h, (void**)&h->u, 3 * N * rs))) return bail(rc);
    if ((rc = dalloc(h, (void**)&h->flags, N))) return bail(rc);
    if ((rc = dalloc(h, (void**)&h->partials, h->n_partials * sizeof(double)))) return bail(rc);
    if (cfg->ext_force_field && (rc = dalloc(h, (void**)&h->F, 3 * N * rs))) return bail(rc);
    if (h->mhd) {
        if ((rc = dalloc(h, (void**)&h->e_stat, 3 * N * rs))) return bail(rc);
        if ((rc = dalloc(h, (void**)&h->b_stat, 3 * N * rs))) return bail(rc);
        if ((rc = dalloc(h, (void**)&h->e_dyn, 3 * N * rs))) return bail(rc);
        if ((rc = dalloc(h, (void**)&h->b_dyn, 3 * N * rs))) return bail(rc);
        if ((rc = dalloc(h, &h->fqi, 7 * N * de))) return bail(rc);
        if ((rc = dalloc(h, &h->ei, N * h->Q * de))) return bail(rc);
        if ((rc = dalloc(h, (void**)&h->q, N * rs))) return bail(rc);
        if ((rc = dalloc(h, (void**)&h->qu_lod, h->n_lod * 4 * rs))) return bail(rc);
        const uint64_t n_top = (uint64_t)h->nd * h->nd * h->nd;
        if ((rc = dalloc(h, (void**)&h->src_pq, n_top * 4 * rs))) return bail(rc);
        if ((rc = dalloc(h, (void**)&h->src_vd, n_top * 4 * rs))) return bail(rc);
        if ((rc = dalloc(h, (void**)&h->src_count, sizeof(uint32_t)))) return bail(rc);
        if (h->lod && (rc = dalloc(h, (void**)&h->lod_seg, (N / h->lod_bx) * 4 * rs))) return bail(rc);
        h->ecr = cfg->ext_subgrid_ecr != 0;
        if (h->ecr) {
            if ((rc = dalloc(h, &h->eti, 7 * N * de))) return bail(rc);
            if ((rc = dalloc(h, (void**)&h->et, N * 4))) return bail(rc);
            if ((rc = dalloc(h, (void**)&h->e_var, 3 * N * 4))) return bail(rc);
        }
    }
    e = ion::launch_fill(h->rho, h->f64(), N, 1.0, h->stream);  // rho buffer starts at 1 (domain.rs:164)
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) { fail(ION_ERR_DEVICE, std::string("initial fill: ") + hipGetErrorString(e)); return bail(ION_ERR_DEVICE); }
    *out = h;
    return ION_OK;
}

void ion_destroy(ion_handle* h) {
    if (!h) return;
    free_all(h);
    delete h;
}

int ion_get_config(const ion_handle* h, ion_config* out) {
    if (!h || !out) return fail(ION_ERR_INVALID, "null argument");
    *out = h->cfg;
    return ION_OK;
}

int ion_initialize(ion_handle* h) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    h->t = 1;  // initialize writes the odd-step layout (mod.rs:216)
    ion::LbmArgs a = lbm_args(h);
    int rc = timed(h, 4, [&] { return ion::launch_initialize(h->Q, h->store, h->mhd, a, h->stream, h->ecr); });
    if (rc) return rc;
    if (h->mhd && (rc = field_update(h))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    h->t = 0;
    h->initialized = true;
    return ION_OK;
}

// The watchdog's check after a step: float64 mass and charge sums, at the
// steps whose counter is a multiple of watch_every.
static int watchdog(ion_handle* h) {
    if (!h->watch_every || h->t % h->watch_every) return ION_OK;
    int rc = ion_conservation(h, &h->watch_mass, &h->watch_charge);
    if (rc) return rc;
    if (std::isfinite(h->watch_mass) && std::isfinite(h->watch_charge)) return ION_OK;
    if (h->watch_first_bad == UINT64_MAX) h->watch_first_bad = h->t;
    return fail(ION_ERR_NONFINITE, "watchdog: non-finite sum at step " + std::to_string(h->t) + " (mass " +
                                       std::to_string(h->watch_mass) + ", charge " +
                                       std::to_string(h->watch_charge) + ")");
}

int ion_do_time_step(ion_handle* h) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    int rc = do_step(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return watchdog(h);
}

int ion_run(ion_handle* h, uint64_t steps) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    int rc;
    if (!h->initialized && (rc = ion_initialize(h))) return rc;
    for (uint64_t i = 0; i < steps; i++) {
        if ((rc = do_step(h))) return rc;
        if ((rc = watchdog(h))) return rc;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return ION_OK;
}

int ion_set_watchdog(ion_handle* h, uint32_t every) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    h->watch_every = every;
    h->watch_first_bad = UINT64_MAX;
    return ION_OK;
}

int ion_watchdog_status(const ion_handle* h, uint64_t* first_nonfinite_step, double* mass, double* charge) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    if (first_nonfinite_step) *first_nonfinite_step = h->watch_first_bad;
    if (mass) *mass = h->watch_mass;
    if (charge) *charge = h->watch_charge;
    return ION_OK;
}

uint64_t ion_get_time_step(const ion_handle* h) { return h ? h->t : 0; }

int ion_update_fields(ion_handle* h) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    ion::LbmArgs a = lbm_args(h);
    int rc = timed(h, 4, [&] { return ion::launch_update_fields(h->Q, h->store, a, h->stream); });
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return ION_OK;
}

int ion_synchronize(ion_handle* h) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return ION_OK;
}

static void* field_ptr(const ion_handle* h, int field, size_t* bytes, uint32_t* esz) {
    const uint64_t N = h->N;
    const size_t de = h->ddf_elem(), rs = h->real();
    void* p = nullptr;
    size_t b = 0;
    uint32_t e = 4;
    switch (field) {
        case ION_FIELD_FI: p = h->fi; b = N * h->Q * de; e = (uint32_t)de; break;
        case ION_FIELD_RHO: p = h->rho; b = N * rs; e = (uint32_t)rs; break;
        case ION_FIELD_U: p = h->u; b = 3 * N * rs; e = (uint32_t)rs; break;
        case ION_FIELD_FLAGS: p = h->flags; b = N; e = 1; break;
        case ION_FIELD_F: p = h->F; b = h->F ? 3 * N * rs : 0; e = (uint32_t)rs; break;
        case ION_FIELD_E_STAT: p = h->e_stat; b = p ? 3 * N * rs : 0; e = (uint32_t)rs; break;
        case ION_FIELD_B_STAT: p = h->b_stat; b = p ? 3 * N * rs : 0; e = (uint32_t)rs; break;
        case ION_FIELD_E_DYN: p = h->e_dyn; b = p ? 3 * N * rs : 0; e = (uint32_t)rs; break;
        case ION_FIELD_B_DYN: p = h->b_dyn; b = p ? 3 * N * rs : 0; e = (uint32_t)rs; break;
        case ION_FIELD_FQI: p = h->fqi; b = p ? 7 * N * de : 0; e = (uint32_t)de; break;
        case ION_FIELD_Q: p = h->q; b = p ? N * rs : 0; e = (uint32_t)rs; break;
        case ION_FIELD_QU_LOD: p = h->qu_lod; b = p ? h->n_lod * 4 * rs : 0; e = (uint32_t)rs; break;
        case ION_FIELD_EI: p = h->ei; b = p ? N * h->Q * de : 0; e = (uint32_t)de; break;
        case ION_FIELD_FIELD_SUMS: p = h->field_raw; b = p ? 6 * N * 4 : 0; break;
        case ION_FIELD_E_VAR: p = h->e_var; b = p ? 3 * N * 4 : 0; break;
        case ION_FIELD_ETI: p = h->eti; b = p ? 7 * N * de : 0; e = (uint32_t)de; break;
        case ION_FIELD_ET: p = h->et; b = p ? N * 4 : 0; break;
        default: b = 0;
    }
    if (bytes) *bytes = b;
    if (esz) *esz = e;
    return b ? p : nullptr;
}

int ion_field_size(const ion_handle* h, int field, size_t* bytes, uint32_t* elem_size) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    if (field < 0 || field >= ION_FIELD_COUNT) return fail(ION_ERR_INVALID, "invalid field id");
    field_ptr(h, field, bytes, elem_size);
    return ION_OK;
}

int ion_write_field(ion_handle* h, int field, const void* src, size_t bytes) {
    if (!h || !src) return fail(ION_ERR_INVALID, "null argument");
    size_t b = 0;
    void* p = field_ptr(h, field, &b, nullptr);
    if (!p) return fail(ION_ERR_INVALID, "field " + std::to_string(field) + " does not exist in this configuration");
    if (bytes != b) return fail(ION_ERR_INVALID, "field " + std::to_string(field) + " has " + std::to_string(b) + " bytes, got " + std::to_string(bytes));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    return ION_OK;
}

int ion_read_field(ion_handle* h, int field, void* dst, size_t bytes) {
    if (!h || !dst) return fail(ION_ERR_INVALID, "null argument");
    size_t b = 0;
    void* p = field_ptr(h, field, &b, nullptr);
    if (!p) return fail(ION_ERR_INVALID, "field " + std::to_string(field) + " does not exist in this configuration");
    if (bytes != b) return fail(ION_ERR_INVALID, "field " + std::to_string(field) + " has " + std::to_string(b) + " bytes, got " + std::to_string(bytes));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(dst, p, bytes, hipMemcpyDeviceToHost));
    return ION_OK;
}

int ion_set_volume_force(ion_handle* h, float fx, float fy, float fz) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    h->fx = fx; h->fy = fy; h->fz = fz;
    h->cfg.f_x = fx; h->cfg.f_y = fy; h->cfg.f_z = fz;
    return ION_OK;
}

int ion_conservation(ion_handle* h, double* mass, double* charge) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    std::vector<double> part(h->n_partials);
    auto sum = [&](const float* x, double* out) -> int {
        hipError_t e = ion::launch_sum_f64(x, h->f64(), h->flags, h->N, h->partials, h->n_partials, h->stream);
        if (e != hipSuccess) return fail(ION_ERR_DEVICE, std::string("sum kernel: ") + hipGetErrorString(e));
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(part.data(), h->partials, part.size() * sizeof(double), hipMemcpyDeviceToHost));
        double s = 0;
        for (double v : part) s += v;
        *out = s;
        return ION_OK;
    };
    int rc;
    if (mass && (rc = sum(h->rho, mass))) return rc;
    if (charge) {
        if (!h->q) *charge = 0;
        else if ((rc = sum(h->q, charge))) return rc;
    }
    return ION_OK;
}

int ion_set_field_mode(ion_handle* h, int mode) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    if (mode != ION_FIELD_EXACT && mode != ION_FIELD_FAST) return fail(ION_ERR_INVALID, "invalid field mode");
    h->field_mode = mode;
    return ION_OK;
}

int ion_set_timing(ion_handle* h, int enable) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    int rc = drain_timing(h);
    if (rc) return rc;
    h->timing = enable != 0;
    h->tm = ion_timing{};
    return ION_OK;
}

int ion_get_timing(ion_handle* h, ion_timing* out) {
    if (!h || !out) return fail(ION_ERR_INVALID, "null argument");
    int rc = drain_timing(h);
    if (rc) return rc;
    *out = h->tm;
    return ION_OK;
}

void ion_surrogate_config_default(ion_surrogate_config* c) {
    c->channels = 16;
    c->seed = 1;
    c->lr = 1e-3f;
    c->lr_head = 1e-4f;
}

int ion_surrogate_create(ion_handle* h, const ion_surrogate_config* cfg) {
    if (!h || !cfg) return fail(ION_ERR_INVALID, "null argument");
    if (!h->mhd) return fail(ION_ERR_INVALID, "the surrogate needs ext_magneto_hydro");
    if (h->f64()) return fail(ION_ERR_UNSUPPORTED, "the surrogate runs on FP32 fields, not on the FP64 instantiation");
    if (h->D < 3) return fail(ION_ERR_UNSUPPORTED, "the surrogate needs mhd_lod_depth >= 3");
    const uint32_t sh = 1u << (h->nd & 31u);
    if (h->cfg.n_x / sh > 1 || h->cfg.n_y / sh > 1 || h->cfg.n_z / sh > 1)
        return fail(ION_ERR_UNSUPPORTED, "the surrogate needs an empty near-field block");
    if (cfg->channels < 8 || cfg->channels % 8 != 0) return fail(ION_ERR_INVALID, "channels must be a positive multiple of 8");
    if (!(cfg->lr >= 0.0f) || !(cfg->lr_head >= 0.0f)) return fail(ION_ERR_INVALID, "learning rates must be >= 0");
    delete h->net;
    h->net = new ion::sur::Net();
    ion::sur::Geom g{h->cfg.n_x, h->cfg.n_y, h->cfg.n_z, h->D, h->N,
                     (uint32_t)(h->n_lod - (uint64_t)h->nd * h->nd * h->nd), h->kc.ke, h->kc.kmu};
    ion::sur::Config c;
    c.c0 = cfg->channels;
    c.seed = cfg->seed;
    c.lr = cfg->lr;
    c.lr_head = cfg->lr_head;
    const int rc = h->net->create(g, c, h->stream);
    if (rc) {
        std::string e = h->net->error();
        delete h->net;
        h->net = nullptr;
        return fail(rc == -2 ? ION_ERR_UNSUPPORTED : ION_ERR_DEVICE, "surrogate: " + e);
    }
    if (!h->field_raw) {
        const int r2 = dalloc(h, (void**)&h->field_raw, (size_t)6 * h->N * sizeof(float));
        if (r2) return r2;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return ION_OK;
}

int ion_compute_fields(ion_handle* h) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    if (!h->mhd) return fail(ION_ERR_INVALID, "fields need ext_magneto_hydro");
    int rc = field_update(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return ION_OK;
}

}  // extern "C"

// Scratch of ion_update_e_b_noalloc: the compacted LOD sources (centre and
// charge, then mean velocity and LOD index: ION_NOALLOC_MAX_SOURCES float4
// each) followed by their count. Every GPU the library runs on holds its own
// copy at its own address; host code uses that address (scratch_on_device),
// never the array's name, which in host code is only a stand-in.
static __device__ __align__(256) float4 scratch[2 * ION_NOALLOC_MAX_SOURCES + 16];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
int scratch_on_device(float4** out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(ION_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(ION_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_ptr[dev] = p;
    }
    *out = (float4*)g_scratch_ptr[dev];
    return ION_OK;
}

}  // namespace

extern "C" {

int ion_update_e_b_noalloc(const ion_config* cfg, const ion_field_buffers* b, int field_mode, void* stream) {
    if (!cfg || !b) return fail(ION_ERR_INVALID, "null argument");
    int rc = validate(*cfg);
    if (rc) return rc;
    if (!cfg->ext_magneto_hydro) return fail(ION_ERR_INVALID, "fields need ext_magneto_hydro");
    if (cfg->float_type == ION_FP64)
        return fail(ION_ERR_UNSUPPORTED, "ion_update_e_b_noalloc takes float fields (FP32, FP16S or FP16C lattices)");
    if (field_mode != ION_FIELD_EXACT && field_mode != ION_FIELD_FAST) return fail(ION_ERR_INVALID, "invalid field mode");
    const uint32_t D = cfg->mhd_lod_depth;
    const uint64_t n_top = 1ull << (3 * D);
    if (n_top > ION_NOALLOC_MAX_SOURCES)
        return fail(ION_ERR_UNSUPPORTED, "mhd_lod_depth " + std::to_string(D) + ": " + std::to_string(n_top) +
                                             " LOD sources, more than the static scratch array holds (" +
                                             std::to_string(ION_NOALLOC_MAX_SOURCES) + ", depth 5)");
    if (!b->flags || !b->qu_lod || !b->e_stat || !b->b_stat || !b->e_dyn || !b->b_dyn)
        return fail(ION_ERR_INVALID, "null device buffer");
    ion::FieldArgs f{};
    const uint32_t nd = 1u << D;
    // near-field block side imax(N/(1<<nd), 1), as in field_update
    const uint32_t sh = 1u << (nd & 31u);
    auto side = [&](uint32_t n) { uint32_t v = n / sh; return v < 1 ? 1u : v; };
    f.dsx = side(cfg->n_x);
    f.dsy = side(cfg->n_y);
    f.dsz = side(cfg->n_z);
    f.near = ((uint64_t)f.dsx * f.dsy * f.dsz > 1) ? 1u : 0u;
    if (f.near && (!b->q || !b->u)) return fail(ION_ERR_INVALID, "the near-field block needs the q and u buffers");
    float4* s = nullptr;
    if ((rc = scratch_on_device(&s))) return rc;
    ion_kernel_constants k;
    ion_kernel_constants_of(cfg, &k);
    const uint32_t hi = (uint32_t)n_lod_own_of(D), lo = (uint32_t)(hi - n_top);
    float4* src_pq = s;
    float4* src_vd = s + ION_NOALLOC_MAX_SOURCES;
    uint32_t* count = (uint32_t*)(s + 2 * ION_NOALLOC_MAX_SOURCES);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = ion::launch_prep_sources(b->qu_lod, false, lo, hi, cfg->n_x, cfg->n_y, cfg->n_z, D, src_pq, src_vd,
                                            count, st);
    if (e != hipSuccess) return fail(ION_ERR_DEVICE, std::string("source compaction: ") + hipGetErrorString(e));
    f.e_stat = b->e_stat;
    f.b_stat = b->b_stat;
    f.e_dyn = b->e_dyn;
    f.b_dyn = b->b_dyn;
    f.flags = b->flags;
    f.q = b->q;
    f.u = b->u;
    f.src_pq = src_pq;
    f.src_vd = src_vd;
    f.src_count = count;
    f.nx = cfg->n_x;
    f.ny = cfg->n_y;
    f.nz = cfg->n_z;
    f.D = D;
    f.N = (uint64_t)cfg->n_x * cfg->n_y * cfg->n_z;
    f.ke = k.ke;
    f.kmu = k.kmu;
    f.raw = nullptr;
    f.f64 = 0u;
    e = ion::launch_update_e_b(f, field_mode == ION_FIELD_FAST, st);
    if (e != hipSuccess) return fail(ION_ERR_DEVICE, std::string("field kernel: ") + hipGetErrorString(e));
    return ION_OK;
}

static int need_net(ion_handle* h) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    if (!h->net) return fail(ION_ERR_INVALID, "no surrogate: call ion_surrogate_create first");
    return ION_OK;
}

int ion_surrogate_infer(ion_handle* h) {
    int rc = need_net(h);
    if (rc || (rc = prep_sources(h))) return rc;
    bool ok = true;
    rc = timed(h, 3, [&] {
        ok = h->net->infer(sur_fields(h)) == 0;
        return hipSuccess;
    });
    if (rc) return rc;
    if (!ok) return fail(ION_ERR_DEVICE, "surrogate: " + h->net->error());
    HIPCHK(hipStreamSynchronize(h->stream));
    return ION_OK;
}

int ion_surrogate_train_step(ion_handle* h, int update, double* rel_e, double* rel_b) {
    int rc = need_net(h);
    if (rc || (rc = prep_sources(h))) return rc;
    double re = 0, rb = 0;
    if (h->net->train_step(sur_fields(h), update != 0, &re, &rb)) return fail(ION_ERR_DEVICE, "surrogate: " + h->net->error());
    HIPCHK(hipStreamSynchronize(h->stream));
    if (update) h->sstats.train_steps += 1;
    if (rel_e) *rel_e = re;
    if (rel_b) *rel_b = rb;
    return ION_OK;
}

int ion_surrogate_residual(ion_handle* h, double* rel_e, double* rel_b) {
    int rc = need_net(h);
    if (rc || (rc = prep_sources(h))) return rc;
    double re = 0, rb = 0;
    if (h->net->residual(sur_fields(h), &re, &rb)) return fail(ION_ERR_DEVICE, "surrogate: " + h->net->error());
    if (rel_e) *rel_e = re;
    if (rel_b) *rel_b = rb;
    return ION_OK;
}

int ion_set_field_solver(ion_handle* h, int solver, uint32_t judge_every, double tolerance) {
    if (!h) return fail(ION_ERR_INVALID, "null handle");
    if (solver != ION_SOLVER_DIRECT && solver != ION_SOLVER_SURROGATE) return fail(ION_ERR_INVALID, "invalid solver");
    if (solver == ION_SOLVER_SURROGATE && !h->net) return fail(ION_ERR_INVALID, "no surrogate: call ion_surrogate_create first");
    if (judge_every == 0) return fail(ION_ERR_INVALID, "judge_every must be >= 1");
    if (!(tolerance >= 0.0)) return fail(ION_ERR_INVALID, "tolerance must be >= 0");  // NaN included
    h->solver = solver;
    h->judge_every = judge_every;
    h->tolerance = tolerance;
    h->field_steps = 0;
    h->fallback_left = 0;
    return ION_OK;
}

int ion_surrogate_set_learning_rate(ion_handle* h, float lr, float lr_head) {
    int rc = need_net(h);
    if (rc) return rc;
    if (!(lr >= 0.0f) || !(lr_head >= 0.0f)) return fail(ION_ERR_INVALID, "learning rates must be >= 0");
    h->net->set_learning_rate(lr, lr_head);
    return ION_OK;
}

int ion_surrogate_get_stats(ion_handle* h, ion_surrogate_stats* out) {
    if (!h || !out) return fail(ION_ERR_INVALID, "null argument");
    *out = h->sstats;
    return ION_OK;
}

int ion_surrogate_param_count(ion_handle* h, uint64_t* n) {
    int rc = need_net(h);
    if (rc) return rc;
    *n = h->net->param_count();
    return ION_OK;
}

int ion_surrogate_read_params(ion_handle* h, int grads, float* dst, uint64_t n) {
    int rc = need_net(h);
    if (rc) return rc;
    if (!dst || n != h->net->param_count()) return fail(ION_ERR_INVALID, "parameter count mismatch");
    if (h->net->copy_params(dst, grads != 0, false)) return fail(ION_ERR_DEVICE, h->net->error());
    return ION_OK;
}

int ion_surrogate_write_params(ion_handle* h, const float* src, uint64_t n) {
    int rc = need_net(h);
    if (rc) return rc;
    if (!src || n != h->net->param_count()) return fail(ION_ERR_INVALID, "parameter count mismatch");
    if (h->net->copy_params(const_cast<float*>(src), false, true)) return fail(ION_ERR_DEVICE, h->net->error());
    HIPCHK(hipStreamSynchronize(h->stream));
    return ION_OK;
}

int ion_device_codec(int float_type, int dir, const void* src, void* dst, uint64_t n) {
    if (float_type != ION_FP16S && float_type != ION_FP16C) return fail(ION_ERR_INVALID, "codec needs FP16S or FP16C");
    if (dir != 0 && dir != 1) return fail(ION_ERR_INVALID, "dir must be 0 (encode) or 1 (decode)");
    if (n == 0) return ION_OK;
    if (!src || !dst) return fail(ION_ERR_INVALID, "null buffer");
    if (n > (1ull << 40)) return fail(ION_ERR_INVALID, "more than 2^40 values");  // keeps n * 4 from wrapping
    const size_t in_b = n * (dir == 0 ? 4 : 2), out_b = n * (dir == 0 ? 2 : 4);
    void *din = nullptr, *dout = nullptr;
    HIPCHK(hipMalloc(&din, in_b));
    hipError_t e = hipMalloc(&dout, out_b);
    if (e == hipSuccess) e = hipMemcpy(din, src, in_b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = ion::launch_codec(float_type, dir, din, dout, n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(dst, dout, out_b, hipMemcpyDeviceToHost);
    (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return fail(ION_ERR_DEVICE, std::string("device codec: ") + hipGetErrorString(e));
    return ION_OK;
}

}  // extern "C"
