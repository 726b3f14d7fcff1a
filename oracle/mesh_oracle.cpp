/*
 * CPU oracle for mesh voxelisation and the static-field precompute.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by the product).
 *
 * Literal restatements of the reference kernels on one domain
 * (DEF_OX = DEF_OY = DEF_OZ = 0, is_halo() false):
 *   voxelize_mesh       src/kernels/sim_kernels.cl:1150-1231, with the host's
 *                       bounds widening and direction choice
 *                       (src/mesh.rs:276-302),
 *   psi_from_mesh       :1234-1251 (loop over every cell, flag test),
 *   static_b_from_mesh  :1253-1276,
 *   static_e_from_mesh  :1278-1299.
 * Operation order as written in the source: dot(a, b) = (ax bx + ay by) + az bz,
 * cross unfused, length(v) = sqrt(dot(v, v)), cb(x) = (x x) x; every
 * operation a separately rounded IEEE float operation (-ffp-contract=off),
 * division and square root correctly rounded. M_PI is a double constant, so
 * psi = (float)((double)psic / (4 pi)).
 *
 * Where the reference reads undefined memory the oracle defines it: a hit
 * with d >= 65536 is counted but not stored and reads as 65535; at most 64
 * stored distances are sorted; with no hits the column is empty.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

struct V3 {
    float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(V3 a) { return sqrtf(dot(a, a)); }
inline float cb(float x) { return x * x * x; }

}  // namespace

extern "C" {

/* tri[9 T]: p0, p1, p2 of each triangle; pmin/pmax: the mesh bounds
 * (Mesh::p_min / p_max), widened here by 2 cells as mesh.rs:283-289 does. */
void orc_voxelize(uint32_t NX, uint32_t NY, uint32_t NZ, uint8_t* flags, float* b_dyn, const float* tri, uint32_t T,
                  const float* pmin, const float* pmax, uint8_t flag, float mx, float my, float mz) {
    const float x0 = pmin[0] - 2.0f, y0 = pmin[1] - 2.0f, z0 = pmin[2] - 2.0f;
    const float x1 = pmax[0] + 2.0f, y1 = pmax[1] + 2.0f, z1 = pmax[2] + 2.0f;
    const float c0 = (y1 - y0) * (z1 - z0), c1 = (z1 - z0) * (x1 - x0), c2 = (x1 - x0) * (y1 - y0);
    const uint32_t direction = (c0 < c1 && c0 < c2) ? 0u : (c1 < c2 ? 1u : 2u);
    const uint32_t A = direction == 0u ? NY * NZ : (direction == 1u ? NZ * NX : NX * NY);
    const uint64_t N = (uint64_t)NX * NY * NZ;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ai = 0; ai < (int64_t)A; ai++) {
        const uint32_t a = (uint32_t)ai;
        uint32_t x, y, z;
        if (direction == 0u) { x = (uint32_t)clampi((int)x0, 0, (int)NX - 1); y = a % NY; z = a / NY; }
        else if (direction == 1u) { x = a / NZ; y = (uint32_t)clampi((int)y0, 0, (int)NY - 1); z = a % NZ; }
        else { x = a % NX; y = a / NX; z = (uint32_t)clampi((int)z0, 0, (int)NZ - 1); }
        // position(xyz) + offset (:131-133, :1169-1170)
        const V3 offset{0.5f * (float)(int)NX - 0.5f, 0.5f * (float)(int)NY - 0.5f, 0.5f * (float)(int)NZ - 0.5f};
        const V3 pos{(float)x + 0.5f - 0.5f * (float)NX, (float)y + 0.5f - 0.5f * (float)NY,
                     (float)z + 0.5f - 0.5f * (float)NZ};
        const V3 r_origin{pos.x + offset.x, pos.y + offset.y, pos.z + offset.z};
        const V3 r_dir{(float)(direction == 0u), (float)(direction == 1u), (float)(direction == 2u)};
        const bool condition =
            direction == 0u ? (r_origin.y < y0 || r_origin.z < z0 || r_origin.y >= y1 || r_origin.z >= z1)
            : direction == 1u ? (r_origin.x < x0 || r_origin.z < z0 || r_origin.x >= x1 || r_origin.z >= z1)
                              : (r_origin.x < x0 || r_origin.y < y0 || r_origin.x >= x1 || r_origin.y >= y1);
        if (condition) continue;
        uint32_t intersections = 0u, intersections_check = 0u;
        uint16_t distances[64];
        for (int k = 0; k < 64; k++) distances[k] = 65535u;
        for (uint32_t i = 0; i < T; i++) {
            const float* p = tri + 9 * (size_t)i;
            const V3 p0i{p[0], p[1], p[2]}, p1i{p[3], p[4], p[5]}, p2i{p[6], p[7], p[8]};
            const V3 u = sub(p1i, p0i), v = sub(p2i, p0i), w = sub(r_origin, p0i);
            const V3 h = cross(r_dir, v), q = cross(w, u);
            const float f = 1.0f / dot(u, h), s = f * dot(w, h), t = f * dot(r_dir, q), d = f * dot(v, q);
            if (s >= 0.0f && s < 1.0f && t >= 0.0f && s + t < 1.0f) {
                if (d > 0.0f) {
                    if (intersections < 64u && d < 65536.0f) distances[intersections] = (uint16_t)d;
                    intersections++;
                } else {
                    intersections_check++;
                }
            }
        }
        const int m = (int)std::min(intersections, 64u);
        for (int i = 1; i < m; i++) {  // insertion sort
            const uint16_t tv = distances[i];
            int j = i - 1;
            while (j >= 0 && distances[j] > tv) {
                distances[j + 1] = distances[j];
                j--;
            }
            distances[j + 1] = tv;
        }
        bool inside = (intersections % 2u) && (intersections_check % 2u);
        uint32_t intersection = (intersections % 2u) != (intersections_check % 2u);
        const uint32_t h0 = direction == 0u ? x : direction == 1u ? y : z;
        const uint32_t hmax = direction == 0u ? (uint32_t)clampi((int)x1, 0, (int)NX)
                              : direction == 1u ? (uint32_t)clampi((int)y1, 0, (int)NY)
                                                : (uint32_t)clampi((int)z1, 0, (int)NZ);
        const uint32_t hmesh = intersections ? h0 + (uint32_t)distances[std::min(intersections - 1u, 63u)] : 0u;
        for (uint32_t hh = h0; hh < hmax; hh++) {
            while (intersection < intersections && hh > h0 + (uint32_t)distances[std::min(intersection, 63u)]) {
                inside = !inside;
                intersection++;
            }
            inside = inside && (intersection < intersections && hh < hmesh);
            const uint64_t n = (uint64_t)(direction == 0u ? hh : x) +
                               ((uint64_t)(direction == 1u ? hh : y) + (uint64_t)(direction == 2u ? hh : z) * NY) * NX;
            if (inside) {
                flags[n] = (uint8_t)((flags[n] & ~0x1F) | flag);  // runtime TYPE_BO = 0x1F
                if (b_dyn) {
                    if (flag & 0x10) {
                        b_dyn[n] = mx;
                        b_dyn[N + n] = my;
                        b_dyn[2 * N + n] = mz;
                    } else if ((flag & 0x08) || (flag & 0x04)) {
                        b_dyn[n] = mx;
                    }
                }
            }
        }
    }
}

/* psi_from_mesh over the padded lattice (psi[(NX+2)(NY+2)(NZ+2)]) followed
 * by static_b_from_mesh into b_stat; M is B_dyn (magnetisation). */
void orc_precompute_b(uint32_t NX, uint32_t NY, uint32_t NZ, const uint8_t* flags, const float* M, float* psi,
                      float* B, float kmu0) {
    const uint64_t N = (uint64_t)NX * NY * NZ;
    const uint32_t PX = NX + 2u, PY = NY + 2u;
    const uint64_t Np = (uint64_t)PX * PY * (NZ + 2u);
#pragma omp parallel for schedule(static)
    for (int64_t ni = 0; ni < (int64_t)Np; ni++) {
        const uint32_t n = (uint32_t)ni;
        const uint32_t t = n % (PX * PY);
        const V3 c{(float)(t % PX), (float)(t / PX), (float)(n / (PX * PY))};
        float psic = 0.0f;
        for (uint64_t i = 0; i < N; i++) {
            if (flags[i] & 0x10) {
                const uint32_t ti = (uint32_t)(i % ((uint64_t)NX * NY));
                const V3 ci{(float)(ti % NX + 1u), (float)(ti / NX + 1u), (float)((uint32_t)(i / ((uint64_t)NX * NY)) + 1u)};
                const V3 cdiff = sub(c, ci);
                const float l = length(cdiff);
                if (!(l == 0.0f)) {
                    const V3 mag{M[i], M[i + N], M[i + 2 * N]};
                    psic += dot(cdiff, mag) / cb(l);
                }
            }
        }
        psi[n] = (float)((double)psic / (4.0 * M_PI));
    }
#pragma omp parallel for schedule(static)
    for (int64_t ni = 0; ni < (int64_t)N; ni++) {
        const uint64_t n = (uint64_t)ni;
        if ((flags[n] & 0x01) == 0x01) continue;
        const uint32_t t = (uint32_t)(n % ((uint64_t)NX * NY));
        const uint32_t x = t % NX + 1u, y = t / NX + 1u, z = (uint32_t)(n / ((uint64_t)NX * NY)) + 1u;
        const uint32_t c = x + (y + z * PY) * PX;  // nabla (:1253-1263)
        const uint32_t yo = PX, zo = PX * PY;
        const float gx = (psi[c + 1] - psi[c - 1]) / 2.0f;
        const float gy = (psi[c + yo] - psi[c - yo]) / 2.0f;
        const float gz = (psi[c + zo] - psi[c - zo]) / 2.0f;
        const float k = -kmu0;
        B[n] += k * gx;
        B[n + N] += k * gy;
        B[n + 2 * N] += k * gz;
    }
}

/* static_e_from_mesh into e_stat; C is B_dyn (charge in its x block). */
void orc_precompute_e(uint32_t NX, uint32_t NY, uint32_t NZ, const uint8_t* flags, const float* C, float* E, float ke) {
    const uint64_t N = (uint64_t)NX * NY * NZ;
#pragma omp parallel for schedule(static)
    for (int64_t ni = 0; ni < (int64_t)N; ni++) {
        const uint64_t n = (uint64_t)ni;
        if ((flags[n] & 0x01) == 0x01) continue;
        const uint32_t t = (uint32_t)(n % ((uint64_t)NX * NY));
        const V3 c{(float)(t % NX), (float)(t / NX), (float)((uint32_t)(n / ((uint64_t)NX * NY)))};
        V3 Ec{0.0f, 0.0f, 0.0f};
        for (uint64_t i = 0; i < N; i++) {
            if ((flags[i] & 0x08) || (flags[i] & 0x04)) {
                const uint32_t ti = (uint32_t)(i % ((uint64_t)NX * NY));
                const V3 ci{(float)(ti % NX), (float)(ti / NX), (float)((uint32_t)(i / ((uint64_t)NX * NY)))};
                const V3 cdiff = sub(c, ci);
                const float l = length(cdiff);
                if (!(l == 0.0f)) {
                    const float q = C[i];
                    const float c3 = cb(l);
                    Ec.x += cdiff.x * q / c3;
                    Ec.y += cdiff.y * q / c3;
                    Ec.z += cdiff.z * q / c3;
                }
            }
        }
        E[n] += Ec.x * ke;
        E[n + N] += Ec.y * ke;
        E[n + 2 * N] += Ec.z * ke;
    }
}

}  // extern "C"
