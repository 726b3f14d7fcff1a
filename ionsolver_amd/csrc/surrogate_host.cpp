// Host side of the learned E/B field surrogate: the 4-level linear 3D conv
// neural operator on the LOD grid, its training step (forward, loss,
// backward, Adam) and inference, all on the GPU through the kernels of
// surrogate.hip.
//
// Network (all convolutions 3^3, zero padding, no bias, no activation; the
// operator it approximates is linear in the LOD charges and currents):
//   x   = (q / s_q, J / s_J) of the top-level LODs, as bf16 hi+lo (8 ch)  grid g0 = nd^3
//   e0  = conv(x)            C0   g0        d2 = conv([up(e3), e2])  C2   g2
//   e1  = conv(pool(e0))     C1   g1        d1 = conv([up(d2), e1])  C1   g1
//   e2  = conv(pool(e1))     C2   g2        d0 = conv([up(d1), e0])  C0   g0
//   e3  = conv(pool(e2))     C3   g3
//   y   = subpixel conv([d0, x]): 6 field components for every cell of the
//         LOD block
// with g_l = nd / 2^l, C_l = c0 (1, 2, 3, 4), plus the global path: the far
// field of each block (all sources outside its 3^3 block neighbourhood),
// summed directly at a 3^3 lattice of points per block and interpolated
// tri-quadratically onto the cells. E = E_stat + KE (s y_E / gain + far_E),
// B likewise with KMU. The head starts at the exact near field (the 26
// neighbouring LOD sources, the own block excluded, as in
// update_e_b_dynamic) and zero weights on the U-Net features; training
// fits the U-Net and the head to what the near + far split leaves.
#include "surrogate_host.hpp"

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "surrogate.hpp"

namespace ion {
namespace sur {

namespace {
inline int pad32(int v) { return (v + 31) / 32 * 32; }
}  // namespace

#define SCHK(x)                                                                           \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            err_ = std::string(#x) + ": " + hipGetErrorString(e_);                        \
            return -3;                                                                    \
        }                                                                                 \
    } while (0)

Net::~Net() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : allocs_) (void)hipFree(p);
}

int Net::alloc(void** p, size_t bytes) {
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        err_ = "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e);
        return -3;
    }
    allocs_.push_back(*p);
    e = hipMemsetAsync(*p, 0, bytes, stream_);
    if (e != hipSuccess) {
        err_ = std::string("hipMemsetAsync: ") + hipGetErrorString(e);
        return -3;
    }
    return 0;
}

int Net::add_layer(int ca, int cb, int cout, int g, Layer* L) {
    L->ca = ca;
    L->cb = cb;
    L->cout = cout;
    L->g = g;
    L->vox = g * g * g;
    L->kpad = pad32(27 * (ca + cb));
    L->cout_pad = pad32(cout);
    L->vox_pad = pad32(L->vox);
    const size_t nw = (size_t)cout * L->kpad;
    int rc;  // inference state; the training state comes with the first training step
    if ((rc = alloc((void**)&L->w32, nw * 4)) || (rc = alloc((void**)&L->w, nw * 2)) ||
        (rc = alloc((void**)&L->wt, (size_t)L->kpad * L->cout_pad * 2)) ||
        (rc = alloc((void**)&L->col, (size_t)L->vox * L->kpad * 2)) ||
        (rc = alloc((void**)&L->out, (size_t)L->vox * cout * 2)))
        return rc;
    return 0;
}

// Gradients, Adam moments and backward scratch, allocated at the first
// training step (an inference-only surrogate never holds them).
int Net::ensure_training_state() {
    if (trainable_) return 0;
    int rc;
    for (int i = 0; i < NLAYER; i++) {
        Layer* L = &layers_[i];
        const size_t nw = (size_t)L->cout * L->kpad;
        if ((rc = alloc((void**)&L->g32, nw * 4)) || (rc = alloc((void**)&L->m, nw * 4)) ||
            (rc = alloc((void**)&L->v, nw * 4)) || (rc = alloc((void**)&L->colT, (size_t)L->kpad * L->vox_pad * 2)) ||
            (rc = alloc((void**)&L->dout32, (size_t)L->vox * L->cout * 4)) ||
            (rc = alloc((void**)&L->dout, (size_t)L->vox * L->cout_pad * 2)) ||
            (rc = alloc((void**)&L->doutT, (size_t)L->cout_pad * L->vox_pad * 2)) ||
            (rc = alloc((void**)&L->dcol, (size_t)L->vox * L->kpad * 4)))
            return rc;
    }
    const int g0 = (int)nd_, C0 = cfg_.c0;
    const int C[4] = {C0, 2 * C0, 3 * C0, 4 * C0};
    for (int l = 0; l < 3; l++) {
        const int g = g0 >> l, h = g / 2;
        if ((rc = alloc((void**)&dpool_[l], (size_t)h * h * h * C[l] * 4)) ||
            (rc = alloc((void**)&dup_[l], (size_t)g * g * g * C[l + 1] * 4)) ||
            (rc = alloc((void**)&de_[l], (size_t)g * g * g * C[l] * 4)))
            return rc;
    }
    if ((rc = alloc((void**)&de_[3], (size_t)(g0 / 8) * (g0 / 8) * (g0 / 8) * C[3] * 4))) return rc;
    const size_t nh = (size_t)nout_ * hk_;
    if ((rc = alloc((void**)&hg32_, nh * 4)) || (rc = alloc((void**)&hm_, nh * 4)) || (rc = alloc((void**)&hv_, nh * 4)) ||
        (rc = alloc((void**)&hcolT_, (size_t)hk_ * nblk_pad_ * 2)) || (rc = alloc((void**)&hdcol_, (size_t)nblk_ * hk_ * 4)) ||
        (rc = alloc((void**)&dY_, (size_t)nblk_ * nout_pad_ * 2)) ||
        (rc = alloc((void**)&dYt_, (size_t)nout_pad_ * nblk_pad_ * 2)))
        return rc;
    trainable_ = true;
    return 0;
}

int Net::create(const Geom& geo, const Config& cfg, hipStream_t stream) {
    geo_ = geo;
    cfg_ = cfg;
    stream_ = stream;
    nd_ = 1u << geo.D;
    if (geo.D < 3) {
        err_ = "the surrogate needs mhd_lod_depth >= 3 (four LOD-grid levels)";
        return -2;
    }
    const int C0 = cfg.c0, C[4] = {C0, 2 * C0, 3 * C0, 4 * C0};
    const int g0 = (int)nd_;
    int rc;
    // encoder / decoder layers
    if ((rc = add_layer(8, 0, C[0], g0, &layers_[E0])) || (rc = add_layer(C[0], 0, C[1], g0 / 2, &layers_[E1])) ||
        (rc = add_layer(C[1], 0, C[2], g0 / 4, &layers_[E2])) || (rc = add_layer(C[2], 0, C[3], g0 / 8, &layers_[E3])) ||
        (rc = add_layer(C[3], C[2], C[2], g0 / 4, &layers_[D2])) ||
        (rc = add_layer(C[2], C[1], C[1], g0 / 2, &layers_[D1])) ||
        (rc = add_layer(C[1], C[0], C[0], g0, &layers_[D0])))
        return rc;
    // pooled / upsampled activations and their gradients
    if ((rc = alloc((void**)&splitk_ws_, SPLITK_CAP * sizeof(float)))) return rc;
    if ((rc = alloc((void**)&x_, (size_t)g0 * g0 * g0 * 8 * 2)) || (rc = alloc((void**)&scale_, 256)) ||
        (rc = alloc((void**)&far_, (size_t)g0 * g0 * g0 * FAR_PTS * 6 * 4)))
        return rc;
    for (int l = 0; l < 3; l++) {
        const int g = g0 >> l, h = g / 2, c = C[l];
        // pooled input of encoder level l + 1, upsampled input of decoder level l
        if ((rc = alloc((void**)&pool_[l], (size_t)h * h * h * c * 2)) ||
            (rc = alloc((void**)&up_[l], (size_t)g * g * g * C[l + 1] * 2)))
            return rc;
    }
    // head
    cpb_ = (geo.nx / nd_) * (geo.ny / nd_) * (geo.nz / nd_);
    gain_ = (float)std::pow((double)cpb_, 2.0 / 3.0);
    nout_ = 6 * cpb_;
    nout_pad_ = pad32((int)nout_);
    nblk_ = nd_ * nd_ * nd_;
    nblk_pad_ = pad32((int)nblk_);
    hcin_ = C0 + 8;
    hk_ = pad32(27 * hcin_);
    const size_t nh = (size_t)nout_ * hk_;
    if ((rc = alloc((void**)&hw32_, nh * 4)) || (rc = alloc((void**)&hw_, nh * 2)) ||
        (rc = alloc((void**)&hwt_, (size_t)hk_ * nout_pad_ * 2)) || (rc = alloc((void**)&hcol_, (size_t)nblk_ * hk_ * 2)))
        return rc;
    ncell_ = geo.ncell;
    n_partials_ = 1024;
    if ((rc = alloc((void**)&partials_, n_partials_ * 4 * sizeof(double))) || (rc = alloc((void**)&norms_, 256)))
        return rc;
    return init_weights();
}

// Network-unit output at every cell, for residuals and training.
int Net::ensure_output_buffer() {
    if (y_) return 0;
    return alloc((void**)&y_, (size_t)6 * ncell_ * 4);
}

// Exact near field of the 26 neighbouring LOD sources for the cell at offset
// o of its block: r = o - (delta + 1/2) b, K = gain r / |r|^3 (E from q, B
// from J x K).
void Net::near_field_head(std::vector<float>& w) const {
    const uint32_t bx = geo_.nx / nd_, by = geo_.ny / nd_, bz = geo_.nz / nd_;
    const int c0 = cfg_.c0;
    for (uint32_t o = 0; o < cpb_; o++) {
        const double ox = o % bx, oy = (o / bx) % by, oz = o / (bx * by);
        for (int d = 0; d < 27; d++) {
            const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
            if (dx == 0 && dy == 0 && dz == 0) continue;
            const double rx = ox - (dx + 0.5) * bx, ry = oy - (dy + 0.5) * by, rz = oz - (dz + 0.5) * bz;
            const double r = std::sqrt(rx * rx + ry * ry + rz * rz);
            const double r3 = r * r * r / gain_;
            const double K[3] = {rx / r3, ry / r3, rz / r3};
            for (int part = 0; part < 2; part++) {  // hi and lo input channels
                const int base = d * hcin_ + c0 + 4 * part;
                auto W = [&](int comp, int ch) -> float& { return w[(size_t)(comp * cpb_ + o) * hk_ + base + ch]; };
                for (int c = 0; c < 3; c++) W(c, 0) = (float)K[c];  // E from q
                W(3, 2) = (float)K[2];  W(3, 3) = (float)-K[1];      // Bx = Jy Kz - Jz Ky
                W(4, 3) = (float)K[0];  W(4, 1) = (float)-K[2];      // By = Jz Kx - Jx Kz
                W(5, 1) = (float)K[1];  W(5, 2) = (float)-K[0];      // Bz = Jx Ky - Jy Kx
            }
        }
    }
}

int Net::init_weights() {
    std::mt19937_64 rng(cfg_.seed);
    for (int i = 0; i < NLAYER; i++) {
        Layer& L = layers_[i];
        const int cin = L.ca + L.cb;
        std::normal_distribution<float> nd(0.0f, 1.0f / std::sqrt(27.0f * cin));
        std::vector<float> w((size_t)L.cout * L.kpad, 0.0f);
        for (int co = 0; co < L.cout; co++)
            for (int k = 0; k < 27 * cin; k++) w[(size_t)co * L.kpad + k] = nd(rng);
        SCHK(hipMemcpyAsync(L.w32, w.data(), w.size() * 4, hipMemcpyHostToDevice, stream_));
        SCHK(hipStreamSynchronize(stream_));
    }
    // head: exact near field on the source channels, zero on the U-Net
    // features (their gradient is not zero, so the U-Net starts learning as
    // soon as the head's feature columns move)
    std::vector<float> hw((size_t)nout_ * hk_, 0.0f);
    near_field_head(hw);
    SCHK(hipMemcpyAsync(hw32_, hw.data(), hw.size() * 4, hipMemcpyHostToDevice, stream_));
    SCHK(hipStreamSynchronize(stream_));
    step_ = 0;
    return refresh_bf16();
}

int Net::refresh_bf16() {
    for (int i = 0; i < NLAYER; i++) {
        Layer& L = layers_[i];
        SCHK(to_bf16(L.w32, L.cout, L.kpad, L.kpad, L.w, 0, 0, stream_));
        SCHK(to_bf16(L.w32, L.cout, L.kpad, L.kpad, L.wt, 1, L.cout_pad, stream_));
    }
    SCHK(to_bf16(hw32_, (int)nout_, hk_, hk_, hw_, 0, 0, stream_));
    SCHK(to_bf16(hw32_, (int)nout_, hk_, hk_, hwt_, 1, nout_pad_, stream_));
    return 0;
}

int Net::conv_fwd(Layer& L, const __bf16* a, const __bf16* b) {
    SCHK(im2col(a, L.ca, b, L.cb, L.g, L.kpad, L.col, 0, 0, stream_));
    GemmOut o{};
    o.bf16 = L.out;
    o.ldc = L.cout;
    o.alpha = 1.0f;
    o.splitk_ws = splitk_ws_;
    o.splitk_cap = SPLITK_CAP;
    SCHK(gemm_nt(EPI_BF16, L.col, L.w, L.vox, L.cout, L.kpad, L.kpad, L.kpad, o, stream_));
    return 0;
}

int Net::forward_features(const float4* qu) {
    const int g0 = (int)nd_, C0 = cfg_.c0;
    const int C[4] = {C0, 2 * C0, 3 * C0, 4 * C0};
    SCHK(input(qu, nblk_, geo_.lo, scale_, x_, stream_));
    int rc;
    if ((rc = conv_fwd(layers_[E0], x_, nullptr))) return rc;
    for (int l = 0; l < 3; l++) {
        SCHK(pool(layers_[E0 + l].out, g0 >> l, C[l], pool_[l], stream_));
        if ((rc = conv_fwd(layers_[E0 + l + 1], pool_[l], nullptr))) return rc;
    }
    // decoder: level 2 (from e3), level 1 (from d2), level 0 (from d1)
    SCHK(up(layers_[E3].out, g0 >> 3, C[3], up_[2], stream_));
    if ((rc = conv_fwd(layers_[D2], up_[2], layers_[E2].out))) return rc;
    SCHK(up(layers_[D2].out, g0 >> 2, C[2], up_[1], stream_));
    if ((rc = conv_fwd(layers_[D1], up_[1], layers_[E1].out))) return rc;
    SCHK(up(layers_[D1].out, g0 >> 1, C[1], up_[0], stream_));
    if ((rc = conv_fwd(layers_[D0], up_[0], layers_[E0].out))) return rc;
    SCHK(im2col(layers_[D0].out, C0, x_, 8, g0, hk_, hcol_, 0, 0, stream_));
    return 0;
}

GemmOut Net::head_out(const Fields& f, float* fine) const {
    GemmOut o{};
    o.alpha = 1.0f;
    o.fine = fine;
    o.e_dyn = f.e_dyn;
    o.b_dyn = f.b_dyn;
    o.e_stat = f.e_stat;
    o.b_stat = f.b_stat;
    o.flags = f.flags;
    o.in_scale = scale_;
    o.scale_e = geo_.ke / gain_;
    o.scale_b = geo_.kmu / gain_;
    o.ke = geo_.ke;
    o.kmu = geo_.kmu;
    o.gain = gain_;
    o.far = far_;
    o.ncell = ncell_;
    o.nx = geo_.nx;
    o.ny = geo_.ny;
    o.nz = geo_.nz;
    o.nd = nd_;
    o.bx = geo_.nx / nd_;
    o.by = geo_.ny / nd_;
    o.bz = geo_.nz / nd_;
    o.cpb = cpb_;
    return o;
}

// Far-field samples of every block (interpolated onto the cells by the
// head GEMM's epilogue).
int Net::far_field(const Fields& f) {
    if (f.src_pq == nullptr || f.src_vd == nullptr || f.src_count == nullptr) {
        err_ = "no LOD sources for the far field";
        return -2;
    }
    FarArgs fa{f.src_pq, f.src_vd, f.src_count, far_, nd_, geo_.D, nblk_, geo_.nx / nd_, geo_.ny / nd_, geo_.nz / nd_};
    SCHK(far_samples(fa, stream_));
    return 0;
}

int Net::infer(const Fields& f) {
    int rc;
    if ((rc = forward_features(f.qu)) || (rc = far_field(f))) return rc;
    SCHK(gemm_nt(EPI_HEAD, hcol_, hw_, (int)nblk_, (int)nout_, hk_, hk_, hk_, head_out(f, nullptr), stream_));
    return 0;
}

int Net::infer_raw(const Fields& f) {
    int rc;
    if ((rc = forward_features(f.qu)) || (rc = far_field(f))) return rc;
    GemmOut o = head_out(f, nullptr);
    o.raw_out = 1;
    SCHK(gemm_nt(EPI_HEAD, hcol_, hw_, (int)nblk_, (int)nout_, hk_, hk_, hk_, o, stream_));
    return 0;
}

int Net::loss(const Fields& f, double* rel_e, double* rel_b) {
    if (f.target == nullptr) {
        err_ = "no direct field to compare with";
        return -2;
    }
    SCHK(loss_sums(y_, f.target, f.flags, ncell_, scale_, gain_, partials_, (int)n_partials_, stream_));
    std::vector<double> p(n_partials_ * 4);
    SCHK(hipMemcpyAsync(p.data(), partials_, p.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    SCHK(hipStreamSynchronize(stream_));
    double s[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < n_partials_; i++)
        for (int k = 0; k < 4; k++) s[k] += p[i * 4 + k];
    // The loss weights each group by 1 / sum t^2. A group whose targets are
    // negligible against the other's (a flow that has barely started moving:
    // J = q v close to 0, so B far below E) would get a weight that
    // overflows the bf16 gradients; it is left out of that step's gradient.
    constexpr double NEGLIGIBLE = 1e-8;
    norm_host_[0] = s[1] > NEGLIGIBLE * s[3] ? (float)s[1] : 0.0f;
    norm_host_[1] = s[3] > NEGLIGIBLE * s[1] ? (float)s[3] : 0.0f;
    if (rel_e) *rel_e = s[1] > 0 ? std::sqrt(s[0] / s[1]) : 0.0;
    if (rel_b) *rel_b = s[3] > 0 ? std::sqrt(s[2] / s[3]) : 0.0;
    return 0;
}

int Net::residual(const Fields& f, double* rel_e, double* rel_b) {
    int rc;
    if ((rc = ensure_output_buffer()) || (rc = forward_features(f.qu)) || (rc = far_field(f))) return rc;
    SCHK(gemm_nt(EPI_HEAD, hcol_, hw_, (int)nblk_, (int)nout_, hk_, hk_, hk_, head_out(f, y_), stream_));
    return loss(f, rel_e, rel_b);
}

// Backward of one conv layer from L.dout32: weight gradient into L.g32, input
// gradients into da / db (fp32, may be null).
int Net::conv_bwd(Layer& L, const __bf16* a, const __bf16* b, float* da, float* db) {
    SCHK(to_bf16(L.dout32, L.vox, L.cout, L.cout_pad, L.dout, 0, 0, stream_));
    SCHK(to_bf16(L.dout32, L.vox, L.cout, L.cout, L.doutT, 1, L.vox_pad, stream_));
    SCHK(im2col(a, L.ca, b, L.cb, L.g, L.kpad, L.colT, 1, L.vox_pad, stream_));
    GemmOut gw{};
    gw.f32 = L.g32;
    gw.ldc = L.kpad;
    gw.alpha = 1.0f;
    SCHK(gemm_nt(EPI_F32, L.doutT, L.colT, L.cout, L.kpad, L.vox_pad, L.vox_pad, L.vox_pad, gw, stream_));
    if (da || db) {
        GemmOut gd{};
        gd.f32 = L.dcol;
        gd.ldc = L.kpad;
        gd.alpha = 1.0f;
        SCHK(gemm_nt(EPI_F32, L.dout, L.wt, L.vox, L.kpad, L.cout_pad, L.cout_pad, L.cout_pad, gd, stream_));
        SCHK(col2im(L.dcol, L.g, L.kpad, L.ca, L.cb, da, db, stream_));
    }
    return 0;
}

int Net::backward() {
    const int g0 = (int)nd_, C0 = cfg_.c0;
    const int C[4] = {C0, 2 * C0, 3 * C0, 4 * C0};
    // head: dW_h = dY^T . col, dcol = dY . W_h
    GemmOut gw{};
    gw.f32 = hg32_;
    gw.ldc = hk_;
    gw.alpha = 1.0f;
    SCHK(gemm_nt(EPI_F32, dYt_, hcolT_, (int)nout_, hk_, nblk_pad_, nblk_pad_, nblk_pad_, gw, stream_));
    GemmOut gd{};
    gd.f32 = hdcol_;
    gd.ldc = hk_;
    gd.alpha = 1.0f;
    SCHK(gemm_nt(EPI_F32, dY_, hwt_, (int)nblk_, hk_, nout_pad_, nout_pad_, nout_pad_, gd, stream_));
    SCHK(col2im(hdcol_, g0, hk_, C0, 8, layers_[D0].dout32, nullptr, stream_));
    int rc;
    // d0 <- [up(d1), e0]
    if ((rc = conv_bwd(layers_[D0], up_[0], layers_[E0].out, dup_[0], de_[0]))) return rc;
    SCHK(up_bwd(dup_[0], g0 >> 1, C[1], layers_[D1].dout32, stream_));
    if ((rc = conv_bwd(layers_[D1], up_[1], layers_[E1].out, dup_[1], de_[1]))) return rc;
    SCHK(up_bwd(dup_[1], g0 >> 2, C[2], layers_[D2].dout32, stream_));
    if ((rc = conv_bwd(layers_[D2], up_[2], layers_[E2].out, dup_[2], de_[2]))) return rc;
    SCHK(up_bwd(dup_[2], g0 >> 3, C[3], layers_[E3].dout32, stream_));
    // encoder, deepest first; e_l's gradient = skip part + pooled part
    if ((rc = conv_bwd(layers_[E3], pool_[2], nullptr, dpool_[2], nullptr))) return rc;
    SCHK(pool_bwd(dpool_[2], g0 >> 2, C[2], de_[2], 1, stream_));
    SCHK(hipMemcpyAsync(layers_[E2].dout32, de_[2], (size_t)layers_[E2].vox * C[2] * 4, hipMemcpyDeviceToDevice, stream_));
    if ((rc = conv_bwd(layers_[E2], pool_[1], nullptr, dpool_[1], nullptr))) return rc;
    SCHK(pool_bwd(dpool_[1], g0 >> 1, C[1], de_[1], 1, stream_));
    SCHK(hipMemcpyAsync(layers_[E1].dout32, de_[1], (size_t)layers_[E1].vox * C[1] * 4, hipMemcpyDeviceToDevice, stream_));
    if ((rc = conv_bwd(layers_[E1], pool_[0], nullptr, dpool_[0], nullptr))) return rc;
    SCHK(pool_bwd(dpool_[0], g0, C[0], de_[0], 1, stream_));
    SCHK(hipMemcpyAsync(layers_[E0].dout32, de_[0], (size_t)layers_[E0].vox * C[0] * 4, hipMemcpyDeviceToDevice, stream_));
    if ((rc = conv_bwd(layers_[E0], x_, nullptr, nullptr, nullptr))) return rc;
    return 0;
}

int Net::train_step(const Fields& f, bool update, double* rel_e, double* rel_b) {
    int rc;
    if ((rc = ensure_output_buffer()) || (rc = ensure_training_state()) || (rc = forward_features(f.qu)) ||
        (rc = far_field(f)))
        return rc;
    SCHK(gemm_nt(EPI_HEAD, hcol_, hw_, (int)nblk_, (int)nout_, hk_, hk_, hk_, head_out(f, y_), stream_));
    double re = 0.0, rb = 0.0;
    if ((rc = loss(f, &re, &rb))) return rc;
    if (rel_e) *rel_e = re;
    if (rel_b) *rel_b = rb;
    // a sample whose loss is not finite (a target or an output that overflowed)
    // would poison the Adam moments: no gradient step for it
    if (!std::isfinite(re) || !std::isfinite(rb)) return 0;
    SCHK(hipMemcpyAsync(norms_, norm_host_, 2 * sizeof(float), hipMemcpyHostToDevice, stream_));
    HeadGeom hg{ncell_, geo_.nx, geo_.ny, geo_.nz, nd_, geo_.nx / nd_, geo_.ny / nd_, geo_.nz / nd_, cpb_,
                (uint32_t)nout_pad_, (uint32_t)nblk_pad_};
    SCHK(im2col(layers_[D0].out, cfg_.c0, x_, 8, (int)nd_, hk_, hcolT_, 1, nblk_pad_, stream_));
    SCHK(loss_grad(y_, f.target, f.flags, hg, scale_, gain_, norms_, dY_, dYt_, stream_));
    if ((rc = backward())) return rc;
    if (!update) return 0;
    step_ += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - std::pow(b1, (float)step_), bc2 = 1.0f - std::pow(b2, (float)step_);
    for (int i = 0; i < NLAYER; i++) {
        Layer& L = layers_[i];
        SCHK(adam(L.w32, L.g32, L.m, L.v, (size_t)L.cout * L.kpad, cfg_.lr, b1, b2, eps, bc1, bc2, stream_));
    }
    SCHK(adam_head(hw32_, hg32_, hm_, hv_, (size_t)nout_ * hk_, hk_, hcin_, cfg_.c0, cfg_.lr, cfg_.lr_head, b1, b2, eps,
                   bc1, bc2, stream_));
    return refresh_bf16();
}

size_t Net::param_count() const {
    size_t n = 0;
    for (int i = 0; i < NLAYER; i++) n += (size_t)layers_[i].cout * layers_[i].kpad;
    return n + (size_t)nout_ * hk_;
}

// Flat fp32 view of all parameters (or gradients): layers E0..E3, D2, D1, D0,
// then the head; each [cout][kpad] row-major.
int Net::copy_params(float* host, bool grads, bool to_device) {
    if (grads) {
        const int rc = ensure_training_state();
        if (rc) return rc;
    }
    size_t off = 0;
    auto cp = [&](float* dev, size_t n) -> int {
        if (to_device) SCHK(hipMemcpyAsync(dev, host + off, n * 4, hipMemcpyHostToDevice, stream_));
        else SCHK(hipMemcpyAsync(host + off, dev, n * 4, hipMemcpyDeviceToHost, stream_));
        off += n;
        return 0;
    };
    int rc;
    for (int i = 0; i < NLAYER; i++)
        if ((rc = cp(grads ? layers_[i].g32 : layers_[i].w32, (size_t)layers_[i].cout * layers_[i].kpad))) return rc;
    if ((rc = cp(grads ? hg32_ : hw32_, (size_t)nout_ * hk_))) return rc;
    SCHK(hipStreamSynchronize(stream_));
    if (to_device && !grads) return refresh_bf16();
    return 0;
}

}  // namespace sur
}  // namespace ion
