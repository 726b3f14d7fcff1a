// Mesh import, voxelisation and static-field precompute: the host side of
// src/mesh.rs (STL reader, transforms, voxelise_mesh) and of
// Lbm::precompute_B / precompute_E (src/lbm/mod.rs:284-314). All host
// arithmetic is f32 in the reference's order (Rust does not contract
// multiply-adds; this file is built with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"
#include "ionsolver.h"
#include "mesh.hpp"

namespace ion {
namespace {

int fail(int code, const std::string& msg) { return ion_internal_set_error(code, msg.c_str()); }

#define HIPCHK(x)                                                                                        \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return fail(ION_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct F3 {
    float x, y, z;
};
F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }

// F32_3_3, src/mesh.rs:37-72 and its products :397-421
struct M33 {
    float xx, yx, zx, xy, yy, zy, xz, yz, zz;
};
F3 operator*(const M33& m, F3 v) {
    return {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.yx * v.x + m.yy * v.y + m.yz * v.z,
            m.zx * v.x + m.zy * v.y + m.zz * v.z};
}
M33 operator*(const M33& a, const M33& m) {
    M33 r;
    r.xx = a.xx * m.xx + a.xy * m.yx + a.xz * m.zx;
    r.xy = a.xx * m.xy + a.xy * m.yy + a.xz * m.zy;
    r.xz = a.xx * m.xz + a.xy * m.yz + a.xz * m.zz;
    r.yx = a.yx * m.xx + a.yy * m.yx + a.yz * m.zx;
    r.yy = a.yx * m.xy + a.yy * m.yy + a.yz * m.zy;
    r.yz = a.yx * m.xz + a.yy * m.yz + a.yz * m.zz;
    r.zx = a.zx * m.xx + a.zy * m.yx + a.zz * m.zx;
    r.zy = a.zx * m.xy + a.zy * m.yy + a.zz * m.zy;
    r.zz = a.zx * m.xz + a.zy * m.yz + a.zz * m.zz;
    return r;
}
float sq(float x) { return x * x; }
M33 rotm_around_v(F3 v, float r) {  // mesh.rs:58-68
    const float sr = std::sin(r), cr = std::cos(r);
    M33 m;
    m.xx = sq(v.x) + (1.0f - sq(v.x)) * cr;
    m.xy = v.x * v.y * (1.0f - cr) - v.z * sr;
    m.xz = v.x * v.z * (1.0f - cr) + v.y * sr;
    m.yx = v.x * v.y * (1.0f - cr) + v.z * sr;
    m.yy = sq(v.y) + (1.0f - sq(v.y)) * cr;
    m.yz = v.y * v.z * (1.0f - cr) - v.x * sr;
    m.zx = v.x * v.z * (1.0f - cr) - v.y * sr;
    m.zy = v.y * v.z * (1.0f - cr) + v.x * sr;
    m.zz = sq(v.z) + (1.0f - sq(v.z)) * cr;
    return m;
}
M33 rotation_matrix(float rx, float ry, float rz) {  // mesh.rs:50-56, degrees -> radians as in mesh.rs:237
    const float pi = 3.14159265358979323846f;
    return rotm_around_v({1.0f, 0.0f, 0.0f}, rx * pi / 180.0f) * rotm_around_v({0.0f, 1.0f, 0.0f}, ry * pi / 180.0f) *
           rotm_around_v({0.0f, 0.0f, 1.0f}, rz * pi / 180.0f);
}

// f32::min / f32::max (NaN-ignoring)
float fmn(float a, float b) { return std::fmin(a, b); }
float fmx(float a, float b) { return std::fmax(a, b); }

}  // namespace

struct Mesh {
    uint32_t T = 0;
    F3 center{}, p_min{}, p_max{};
    std::vector<F3> p0, p1, p2;

    void update_bounds() {  // mesh.rs:97-113
        p_min = p0[0];
        p_max = p0[0];
        for (uint32_t i = 0; i < T; i++) {
            const F3 a = p0[i], b = p1[i], c = p2[i];
            p_min.x = fmn(p_min.x, fmn(c.x, fmn(b.x, a.x)));
            p_min.y = fmn(p_min.y, fmn(c.y, fmn(b.y, a.y)));
            p_min.z = fmn(p_min.z, fmn(c.z, fmn(b.z, a.z)));
            p_max.x = fmx(p_max.x, fmx(c.x, fmx(b.x, a.x)));
            p_max.y = fmx(p_max.y, fmx(c.y, fmx(b.y, a.y)));
            p_max.z = fmx(p_max.z, fmx(c.z, fmx(b.z, a.z)));
        }
    }
};

struct MeshSet {
    std::vector<Mesh> meshes;
};

void mesh_set_destroy(MeshSet* m) { delete m; }

namespace {

// Mesh::read_stl_raw, src/mesh.rs:175-217
int read_stl(const char* path, bool reposition, F3 box, F3 center, const M33& rot, float size, Mesh* out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(ION_ERR_IO, std::string("could not open or find file \"") + path + "\"");
    std::vector<uint8_t> b;
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + k);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return fail(ION_ERR_IO, std::string("read error: ") + path);
    uint32_t T = 0;
    if (b.size() >= 84) std::memcpy(&T, b.data() + 80, 4);
    if (!(b.size() >= 84 && T > 0 && (uint64_t)b.size() == 84ull + 50ull * T))
        return fail(ION_ERR_FORMAT, std::string("File \"") + path +
                                        "\" is corrupted or unsupported. Only binary .stl files are supported.");
    Mesh m;
    m.T = T;
    m.center = center;
    m.p0.resize(T);
    m.p1.resize(T);
    m.p2.resize(T);
    const uint8_t* p = b.data() + 84;
    auto rd = [&](F3* v) {
        float c[3];
        std::memcpy(c, p, 12);
        p += 12;
        *v = rot * F3{c[0], c[1], c[2]};
    };
    for (uint32_t i = 0; i < T; i++) {
        p += 12;  // normal
        rd(&m.p0[i]);
        rd(&m.p1[i]);
        rd(&m.p2[i]);
        p += 2;  // attribute bits
    }
    m.update_bounds();
    float scale;
    if (size == 0.0f) {  // get_scale_for_box_fit, mesh.rs:160-162
        scale = fmn(box.x / (m.p_max.x - m.p_min.x),
                    fmn(box.y / (m.p_max.y - m.p_min.y), box.z / (m.p_max.z - m.p_min.z)));
    } else if (size > 0.0f) {  // get_max_size, mesh.rs:154-156
        scale = size / fmx(m.p_max.x - m.p_min.x, fmx(m.p_max.y - m.p_min.y, m.p_max.z - m.p_min.z));
    } else {
        scale = -size;
    }
    const F3 offset = reposition ? -0.5f * (m.p_min + m.p_max) : F3{0.0f, 0.0f, 0.0f};
    for (uint32_t i = 0; i < T; i++) {
        m.p0[i] = center + scale * (offset + m.p0[i]);
        m.p1[i] = center + scale * (offset + m.p1[i]);
        m.p2[i] = center + scale * (offset + m.p2[i]);
    }
    m.update_bounds();
    *out = std::move(m);
    return ION_OK;
}

int add_mesh(ion_handle* h, Mesh&& m, uint32_t* index) {
    HandleView v;
    int rc = ion_internal_view(h, &v);
    if (rc) return rc;
    if (!*v.meshes) *v.meshes = new MeshSet();
    (*v.meshes)->meshes.push_back(std::move(m));
    if (index) *index = (uint32_t)(*v.meshes)->meshes.size() - 1u;
    return ION_OK;
}

int get_mesh(ion_handle* h, uint32_t index, HandleView* v, const Mesh** m) {
    int rc = ion_internal_view(h, v);
    if (rc) return rc;
    if (!*v->meshes || index >= (*v->meshes)->meshes.size())
        return fail(ION_ERR_INVALID, "mesh index " + std::to_string(index) + " out of range");
    *m = &(*v->meshes)->meshes[index];
    return ION_OK;
}

// Sources of one kind compacted on the device; freed by the destructor.
struct SourceList {
    uint32_t* counts = nullptr;  // block offsets + total
    uint32_t* total = nullptr;
    float4* pos = nullptr;
    float4* val = nullptr;
    uint32_t n = 0;
    ~SourceList() {
        (void)hipFree(counts);
        (void)hipFree(total);
        (void)hipFree(pos);
        (void)hipFree(val);
    }
};

int select_sources(const HandleView& v, int mode, SourceList* L) {
    const uint32_t nb = mesh::select_blocks(v.N);
    HIPCHK(hipMalloc((void**)&L->counts, (size_t)(nb + 1) * 4));
    HIPCHK(hipMalloc((void**)&L->total, 8));  // count, magnetisation-range flag
    HIPCHK(hipMemsetAsync(L->total, 0, 8, v.stream));
    HIPCHK(mesh::launch_select_count(mode, v.flags, v.N, L->counts, L->total, v.stream));
    HIPCHK(hipMemcpyAsync(&L->n, L->total, 4, hipMemcpyDeviceToHost, v.stream));
    HIPCHK(hipStreamSynchronize(v.stream));
    const size_t n = L->n ? L->n : 1;
    HIPCHK(hipMalloc((void**)&L->pos, n * sizeof(float4)));
    HIPCHK(hipMalloc((void**)&L->val, n * sizeof(float4)));
    mesh::Sources s{L->pos, L->val, L->total};
    HIPCHK(mesh::launch_select_scatter(mode, v.flags, v.b_dyn, v.N, v.nx, v.ny, L->counts, s, v.stream));
    return ION_OK;
}

int need_mhd(const HandleView& v, const char* what) {
    if (!v.mhd) return fail(ION_ERR_INVALID, std::string(what) + " needs ext_magneto_hydro");
    if (v.f64) return fail(ION_ERR_UNSUPPORTED, std::string(what) + " is not built for the FP64 instantiation");
    return ION_OK;
}

}  // namespace
}  // namespace ion

using namespace ion;

extern "C" {

int ion_import_mesh(ion_handle* h, const char* path, float scale, float ox, float oy, float oz, float rx, float ry,
                    float rz, uint32_t* index) {
    HandleView v;
    int rc = ion_internal_view(h, &v);
    if (rc) return rc;
    if (!path) return fail(ION_ERR_INVALID, "null path");
    // mesh.rs:234-240: scale_lu = len_si_lu(scale); size = -|scale_lu|
    const float scale_lu = scale / v.cfg->units.m;
    Mesh m;
    if ((rc = read_stl(path, false, F3{1.0f, 1.0f, 1.0f}, F3{ox, oy, oz}, rotation_matrix(rx, ry, rz),
                       -std::fabs(scale_lu), &m)))
        return rc;
    return add_mesh(h, std::move(m), index);
}

int ion_import_mesh_reposition(ion_handle* h, const char* path, float cx, float cy, float cz, float rx, float ry,
                               float rz, float size, uint32_t* index) {
    HandleView v;
    int rc = ion_internal_view(h, &v);
    if (rc) return rc;
    if (!path) return fail(ION_ERR_INVALID, "null path");
    const F3 box{(float)v.nx, (float)v.ny, (float)v.nz};  // mesh.rs:249
    Mesh m;
    if ((rc = read_stl(path, true, box, F3{cx, cy, cz}, rotation_matrix(rx, ry, rz), size, &m))) return rc;
    return add_mesh(h, std::move(m), index);
}

int ion_mesh_get(ion_handle* h, uint32_t index, uint32_t* triangles, float* vertices, float* bounds) {
    HandleView v;
    const Mesh* m;
    int rc = get_mesh(h, index, &v, &m);
    if (rc) return rc;
    if (triangles) *triangles = m->T;
    if (vertices)
        for (uint32_t i = 0; i < m->T; i++) {
            const F3 q[3] = {m->p0[i], m->p1[i], m->p2[i]};
            for (int k = 0; k < 3; k++) {
                vertices[9 * (size_t)i + 3 * k + 0] = q[k].x;
                vertices[9 * (size_t)i + 3 * k + 1] = q[k].y;
                vertices[9 * (size_t)i + 3 * k + 2] = q[k].z;
            }
        }
    if (bounds) {
        const float b[9] = {m->p_min.x, m->p_min.y, m->p_min.z, m->p_max.x, m->p_max.y,
                            m->p_max.z, m->center.x, m->center.y, m->center.z};
        std::memcpy(bounds, b, sizeof b);
    }
    return ION_OK;
}

int ion_voxelise_mesh(ion_handle* h, uint32_t index, int model_type, float vx, float vy, float vz) {
    HandleView v;
    const Mesh* m;
    int rc = get_mesh(h, index, &v, &m);
    if (rc) return rc;
    if (v.f64) return fail(ION_ERR_UNSUPPORTED, "voxelisation is not built for the FP64 instantiation");
    uint8_t flag;
    switch (model_type) {  // mesh.rs:309-314
        case ION_MODEL_SOLID: flag = 0x01; break;
        case ION_MODEL_MAGNET: flag = 0x11; break;
        case ION_MODEL_CHARGED: flag = 0x09; break;
        case ION_MODEL_CHARGED_ECR: flag = 0x05; break;
        default: return fail(ION_ERR_INVALID, "invalid model_type");
    }
    mesh::VoxArgs a{};
    // bounds widened by 2 cells (mesh.rs:283-289)
    a.x0 = m->p_min.x - 2.0f;
    a.y0 = m->p_min.y - 2.0f;
    a.z0 = m->p_min.z - 2.0f;
    a.x1 = m->p_max.x + 2.0f;
    a.y1 = m->p_max.y + 2.0f;
    a.z1 = m->p_max.z + 2.0f;
    const float b[6] = {a.x0, a.y0, a.z0, a.x1, a.y1, a.z1};
    for (float x : b)
        if (!std::isfinite(x) || std::fabs(x) > 2.0e9f) return fail(ION_ERR_INVALID, "mesh bounds not finite");
    // direction of the smallest bounding-box cross-section (mesh.rs:293-302)
    const float c0 = (a.y1 - a.y0) * (a.z1 - a.z0), c1 = (a.z1 - a.z0) * (a.x1 - a.x0), c2 = (a.x1 - a.x0) * (a.y1 - a.y0);
    a.direction = (c0 < c1 && c0 < c2) ? 0u : (c1 < c2 ? 1u : 2u);
    a.nx = v.nx;
    a.ny = v.ny;
    a.nz = v.nz;
    a.N = v.N;
    a.A = a.direction == 0u ? v.ny * v.nz : (a.direction == 1u ? v.nx * v.nz : v.nx * v.ny);
    a.T = m->T;
    a.flags = v.flags;
    a.flag = flag;
    a.b_dyn = v.mhd ? v.b_dyn : nullptr;
    const ion_units& u = v.cfg->units;
    if (model_type == ION_MODEL_MAGNET) {  // magnetization_si_lu, units.rs:145-147
        a.mx = vx / (u.a / u.m);
        a.my = vy / (u.a / u.m);
        a.mz = vz / (u.a / u.m);
    } else if (model_type == ION_MODEL_CHARGED || model_type == ION_MODEL_CHARGED_ECR) {  // charge_si_lu, :130-133
        a.mx = vx / (u.a * u.s);
    }
    std::vector<float> tri(9 * (size_t)m->T);
    (void)ion_mesh_get(h, index, nullptr, tri.data(), nullptr);
    float* d_tri = nullptr;
    HIPCHK(hipMalloc((void**)&d_tri, tri.size() * sizeof(float)));
    a.tri = d_tri;
    hipError_t e = hipMemcpyAsync(d_tri, tri.data(), tri.size() * sizeof(float), hipMemcpyHostToDevice, v.stream);
    if (e == hipSuccess) e = mesh::launch_voxelize(a, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    (void)hipFree(d_tri);
    if (e != hipSuccess) return fail(ION_ERR_DEVICE, std::string("voxelize_mesh: ") + hipGetErrorString(e));
    return ION_OK;
}

int ion_precompute_b(ion_handle* h) {
    HandleView v;
    int rc = ion_internal_view(h, &v);
    if (rc || (rc = need_mhd(v, "precompute_B"))) return rc;
    SourceList L;
    if ((rc = select_sources(v, 0, &L))) return rc;
    // psi over the padded lattice; in E_dyn as in the reference
    // (src/lbm/domain.rs:280), or in a scratch buffer where E_dyn is smaller
    const uint64_t Np = (uint64_t)(v.nx + 2) * (v.ny + 2) * (v.nz + 2);
    float* psi = v.e_dyn;
    float* tmp = nullptr;
    if (Np > 3 * v.N) {
        HIPCHK(hipMalloc((void**)&tmp, Np * sizeof(float)));
        psi = tmp;
    }
    mesh::Sources s{L.pos, L.val, L.total};
    hipError_t e = mesh::launch_psi(s, psi, v.nx, v.ny, v.nz, v.stream);
    if (e == hipSuccess) e = mesh::launch_static_b(v.flags, v.b_stat, psi, v.nx, v.ny, v.nz, v.kc->kmu0, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(ION_ERR_DEVICE, std::string("precompute_B: ") + hipGetErrorString(e));
    return ION_OK;
}

int ion_psi_math_check(uint64_t* sqrt_mismatch, uint64_t* div_mismatch, uint64_t* div_samples) {
    if (!sqrt_mismatch || !div_mismatch || !div_samples) return fail(ION_ERR_INVALID, "null argument");
    HIPCHK(mesh::run_psi_math_check(sqrt_mismatch, div_mismatch, div_samples));
    return ION_OK;
}

int ion_precompute_e_ecr(ion_handle* h) {
    HandleView v;
    int rc = ion_internal_view(h, &v);
    if (rc) return rc;
    if (!v.e_var) return fail(ION_ERR_INVALID, "precompute_E_ECR needs ext_magneto_hydro and ext_subgrid_ecr");
    SourceList L;
    if ((rc = select_sources(v, 1, &L))) return rc;
    mesh::Sources s{L.pos, L.val, L.total};
    hipError_t e = mesh::launch_static_e(s, v.flags, v.e_var, v.nx, v.ny, v.nz, v.kc->ke, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    if (e != hipSuccess) return fail(ION_ERR_DEVICE, std::string("precompute_E_ECR: ") + hipGetErrorString(e));
    return ION_OK;
}

int ion_precompute_e(ion_handle* h) {
    HandleView v;
    int rc = ion_internal_view(h, &v);
    if (rc || (rc = need_mhd(v, "precompute_E"))) return rc;
    SourceList L;
    if ((rc = select_sources(v, 1, &L))) return rc;
    mesh::Sources s{L.pos, L.val, L.total};
    hipError_t e = mesh::launch_static_e(s, v.flags, v.e_stat, v.nx, v.ny, v.nz, v.kc->ke, v.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
    if (e != hipSuccess) return fail(ION_ERR_DEVICE, std::string("precompute_E: ") + hipGetErrorString(e));
    return ION_OK;
}

}  // extern "C"
