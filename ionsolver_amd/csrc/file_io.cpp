// Save/load of a simulation in the reference's binary format (FILE_LAYOUT.txt,
// encode/decode in src/file.rs:42-306), little endian:
//   "IonSolver setup\n" | 62-byte config | flags N u8 | rho N f32 | u 3N f32 (SoA)
//   | [MHD] Q N f32 | N_C u32 + 12 N_C bytes | N_M u32 + 20 N_M bytes
// Macroscopic fields only: the DDFs are rebuilt by initialize() after loading.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ionsolver.h"

namespace {

struct Writer {
    std::vector<uint8_t> b;
    void u8(uint8_t v) { b.push_back(v); }
    void u32(uint32_t v) {
        for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i)));
    }
    void f32(float v) {
        uint32_t u;
        std::memcpy(&u, &v, 4);
        u32(u);
    }
    void raw(const void* p, size_t n) {
        const uint8_t* q = (const uint8_t*)p;
        b.insert(b.end(), q, q + n);
    }
};

struct Reader {
    const std::vector<uint8_t>& b;
    size_t pos = 0;
    bool ok = true;
    explicit Reader(const std::vector<uint8_t>& v) : b(v) {}
    bool need(size_t n) {
        if (pos + n > b.size()) ok = false;
        return ok;
    }
    uint8_t u8() { return need(1) ? b[pos++] : 0; }
    uint32_t u32() {
        if (!need(4)) return 0;
        uint32_t v = 0;
        for (int i = 0; i < 4; i++) v |= (uint32_t)b[pos++] << (8 * i);
        return v;
    }
    float f32() {
        const uint32_t u = u32();
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    }
    const uint8_t* take(size_t n) {
        if (!need(n)) return nullptr;
        const uint8_t* p = b.data() + pos;
        pos += n;
        return p;
    }
    bool at_end() const { return pos == b.size(); }
};

const char kHeader[17] = "IonSolver setup\n";

}  // namespace

// The C ABI error string lives in lbm_host.cpp; these entry points report
// through the same channel by calling back into an accessor there.
extern "C" __attribute__((visibility("hidden"))) int ion_internal_set_error(int code, const char* msg);

extern "C" int ion_save(ion_handle* h, const char* path) {
    if (!h || !path) return ion_internal_set_error(ION_ERR_INVALID, "null argument");
    ion_config c;
    int rc = ion_get_config(h, &c);
    if (rc) return rc;
    if (c.float_type == ION_FP64)
        return ion_internal_set_error(ION_ERR_UNSUPPORTED, "the file format has no FP64 float type (FILE_LAYOUT.txt)");
    const uint64_t N = (uint64_t)c.n_x * c.n_y * c.n_z;
    Writer w;
    w.raw(kHeader, 16);
    w.u8((uint8_t)c.velocity_set);
    w.u8((uint8_t)c.relaxation_time);
    w.u8((uint8_t)c.float_type);
    w.f32(c.units.m);
    w.f32(c.units.kg);
    w.f32(c.units.s);
    w.f32(c.units.a);
    w.u32(c.n_x);
    w.u32(c.n_y);
    w.u32(c.n_z);
    w.u8(0);  // fixed domain sizes
    w.u32(c.d_x);
    w.u32(c.d_y);
    w.u32(c.d_z);
    w.f32(c.nu);
    w.f32(c.f_x);
    w.f32(c.f_y);
    w.f32(c.f_z);
    w.u8((uint8_t)((c.ext_equilibrium_boudaries ? 1 : 0) | (c.ext_volume_force ? 2 : 0) | (c.ext_force_field ? 4 : 0) |
                   (c.ext_magneto_hydro ? 8 : 0)));
    w.u8(c.mhd_lod_depth);
    const size_t hdr = w.b.size();
    const bool mhd = c.ext_magneto_hydro != 0;
    w.b.resize(hdr + N * (1 + 4 + 12 + (mhd ? 4 : 0)));
    uint8_t* p = w.b.data() + hdr;
    if ((rc = ion_read_field(h, ION_FIELD_FLAGS, p, N))) return rc;
    p += N;
    if ((rc = ion_read_field(h, ION_FIELD_RHO, p, 4 * N))) return rc;
    p += 4 * N;
    if ((rc = ion_read_field(h, ION_FIELD_U, p, 12 * N))) return rc;
    p += 12 * N;
    if (mhd) {
        if ((rc = ion_read_field(h, ION_FIELD_Q, p, 4 * N))) return rc;
        w.u32(0);  // N_C: no fixed charges
        w.u32(0);  // N_M: no magnets
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return ion_internal_set_error(ION_ERR_IO, (std::string("cannot open for writing: ") + path).c_str());
    const size_t n = std::fwrite(w.b.data(), 1, w.b.size(), f);
    const int cl = std::fclose(f);
    if (n != w.b.size() || cl != 0) return ion_internal_set_error(ION_ERR_IO, (std::string("short write: ") + path).c_str());
    return ION_OK;
}

extern "C" int ion_load(const char* path, ion_config* cfg, ion_handle** out) {
    if (!path || !cfg || !out) return ion_internal_set_error(ION_ERR_INVALID, "null argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return ion_internal_set_error(ION_ERR_IO, (std::string("cannot open: ") + path).c_str());
    std::vector<uint8_t> buf;
    uint8_t tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    Reader r(buf);
    const uint8_t* head = r.take(16);
    if (!head || std::memcmp(head, kHeader, 16) != 0) return ion_internal_set_error(ION_ERR_FORMAT, "Invalid Format!");
    ion_config c = *cfg;
    const uint8_t vs = r.u8(), rt = r.u8(), ft = r.u8();
    if (vs > 3 || rt > 1 || ft > 2) return ion_internal_set_error(ION_ERR_FORMAT, "invalid enum value in header");
    c.velocity_set = vs;
    c.relaxation_time = rt;
    c.float_type = ft;  // FloatType discriminants of src/lbm/types.rs:75-83, as written by encode
    c.units.m = r.f32();
    c.units.kg = r.f32();
    c.units.s = r.f32();
    c.units.a = r.f32();
    c.n_x = r.u32();
    c.n_y = r.u32();
    c.n_z = r.u32();
    r.u8();  // fixed domains (skipped, as the reference does)
    c.d_x = r.u32();
    c.d_y = r.u32();
    c.d_z = r.u32();
    c.nu = r.f32();
    c.f_x = r.f32();
    c.f_y = r.f32();
    c.f_z = r.f32();
    const uint8_t ext = r.u8();
    c.ext_equilibrium_boudaries = (ext & 1) != 0;
    c.ext_volume_force = (ext & 2) != 0;
    c.ext_force_field = (ext & 4) != 0;
    c.ext_magneto_hydro = (ext & 8) != 0;
    c.mhd_lod_depth = r.u8();
    if (!r.ok) return ion_internal_set_error(ION_ERR_FORMAT, "truncated header");
    // the section lengths follow from the lattice: bounded first (n_x * n_y
    // before n_z, so that the product cannot wrap), as ion_create bounds it
    const uint64_t nxy = (uint64_t)c.n_x * c.n_y;
    if (nxy > 0x40000000ull || nxy * c.n_z > 0x40000000ull)
        return ion_internal_set_error(ION_ERR_UNSUPPORTED, "lattice of more than 2^30 cells in the header");
    const uint64_t N = nxy * c.n_z;
    const bool mhd = c.ext_magneto_hydro != 0;
    const uint8_t* flags = r.take(N);
    const uint8_t* rho = r.take(4 * N);
    const uint8_t* u = r.take(12 * N);
    const uint8_t* q = mhd ? r.take(4 * N) : nullptr;
    if (!r.ok) return ion_internal_set_error(ION_ERR_FORMAT, "truncated field sections");
    if (mhd && !r.at_end()) {
        // fixed charges and magnets (read and discarded by the reference)
        const uint32_t nc = r.u32();
        r.take((size_t)nc * 12);
        const uint32_t nm = r.u32();
        r.take((size_t)nm * 20);
        if (!r.ok || !r.at_end())
            return ion_internal_set_error(ION_ERR_FORMAT, "Not all data could be read, file may be corrupted.");
    }
    ion_handle* h = nullptr;
    int rc = ion_create(&c, &h);
    if (rc) return rc;
    if ((rc = ion_write_field(h, ION_FIELD_FLAGS, flags, N)) || (rc = ion_write_field(h, ION_FIELD_RHO, rho, 4 * N)) ||
        (rc = ion_write_field(h, ION_FIELD_U, u, 12 * N)) || (mhd && (rc = ion_write_field(h, ION_FIELD_Q, q, 4 * N)))) {
        ion_destroy(h);
        return rc;
    }
    *cfg = c;
    *out = h;
    return ION_OK;
}
