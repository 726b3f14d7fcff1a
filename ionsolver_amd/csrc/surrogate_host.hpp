// The learned E/B field surrogate network (see surrogate_host.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "surrogate.hpp"

namespace ion {
namespace sur {

struct Geom {
    uint32_t nx, ny, nz, D;
    uint64_t ncell;
    uint32_t lo;   // first LOD index the direct field update reads
    float ke, kmu; // DEF_KE, DEF_KMU
};

struct Config {
    int c0 = 16;          // channels at level 0 (levels: c0, 2c0, 3c0, 4c0)
    uint64_t seed = 1;
    float lr = 1e-3f;     // Adam step of the U-Net layers
    float lr_head = 1e-4f;
};

// Device pointers of the fields the surrogate reads and writes. `target`
// holds the direct kernel's raw source sums (E then B, 6 x N), the training
// and judging reference.
struct Fields {
    const float4* qu;
    const uint8_t* flags;
    const float* e_stat;
    const float* b_stat;
    float* e_dyn;
    float* b_dyn;
    const float* target;
    const float4* src_pq;  // compacted LOD sources of the direct kernel (prep_sources)
    const float4* src_vd;
    const uint32_t* src_count;
};

class Net {
  public:
    enum { E0 = 0, E1, E2, E3, D2, D1, D0, NLAYER };
    ~Net();
    int create(const Geom& geo, const Config& cfg, hipStream_t stream);
    // surrogate field: E_dyn / B_dyn of fluid cells from the LOD sums
    int infer(const Fields& f);
    // The surrogate's raw source sums (the dynamic parts before the KE / KMU
    // scaling) of the fluid cells, written into f.e_dyn / f.b_dyn.
    int infer_raw(const Fields& f);
    // forward + relative L2 errors against the raw direct sums in f.target
    int residual(const Fields& f, double* rel_e, double* rel_b);
    // forward, loss against f.target, backward and (if update) one Adam step
    int train_step(const Fields& f, bool update, double* rel_e, double* rel_b);
    size_t param_count() const;
    void set_learning_rate(float lr, float lr_head) {
        cfg_.lr = lr;
        cfg_.lr_head = lr_head;
    }
    int copy_params(float* host, bool grads, bool to_device);
    int init_weights();
    const std::string& error() const { return err_; }
    uint64_t steps() const { return step_; }

  private:
    struct Layer {
        int ca = 0, cb = 0, cout = 0, g = 0, vox = 0, kpad = 0, cout_pad = 0, vox_pad = 0;
        float *w32 = nullptr, *g32 = nullptr, *m = nullptr, *v = nullptr;
        __bf16 *w = nullptr, *wt = nullptr, *col = nullptr, *colT = nullptr, *out = nullptr, *dout = nullptr,
               *doutT = nullptr;
        float *dout32 = nullptr, *dcol = nullptr;
    };
    int alloc(void** p, size_t bytes);
    int add_layer(int ca, int cb, int cout, int g, Layer* L);
    int refresh_bf16();
    int conv_fwd(Layer& L, const __bf16* a, const __bf16* b);
    int conv_bwd(Layer& L, const __bf16* a, const __bf16* b, float* da, float* db);
    int forward_features(const float4* qu);
    int backward();
    int loss(const Fields& f, double* rel_e, double* rel_b);
    int ensure_output_buffer();
    int ensure_training_state();
    void near_field_head(std::vector<float>& w) const;
    GemmOut head_out(const Fields& f, float* fine) const;
    int far_field(const Fields& f);

    Geom geo_{};
    Config cfg_{};
    hipStream_t stream_ = nullptr;
    std::string err_;
    std::vector<void*> allocs_;
    uint32_t nd_ = 1;
    float gain_ = 1.0f;  // network output = raw field sum * gain / s (gain = block side^2)
    Layer layers_[NLAYER];
    __bf16* x_ = nullptr;
    float* far_ = nullptr;  // far-field samples [block][6][27]
    // split-K partial sums of the U-Net's forward GEMMs (few output tiles, long K)
    static constexpr size_t SPLITK_CAP = (size_t)1 << 20;
    float* splitk_ws_ = nullptr;
    float* scale_ = nullptr;
    __bf16* pool_[3] = {nullptr, nullptr, nullptr};
    float* dpool_[3] = {nullptr, nullptr, nullptr};
    __bf16* up_[3] = {nullptr, nullptr, nullptr};
    float* dup_[3] = {nullptr, nullptr, nullptr};
    float* de_[4] = {nullptr, nullptr, nullptr, nullptr};
    // head
    uint32_t cpb_ = 0, nout_ = 0, nblk_ = 0;
    int nout_pad_ = 0, nblk_pad_ = 0, hcin_ = 0, hk_ = 0;
    float *hw32_ = nullptr, *hg32_ = nullptr, *hm_ = nullptr, *hv_ = nullptr, *hdcol_ = nullptr;
    __bf16 *hw_ = nullptr, *hwt_ = nullptr, *hcol_ = nullptr, *hcolT_ = nullptr;
    // training
    uint64_t ncell_ = 0;
    float* y_ = nullptr;
    __bf16 *dY_ = nullptr, *dYt_ = nullptr;
    double* partials_ = nullptr;
    size_t n_partials_ = 0;
    float* norms_ = nullptr;
    float norm_host_[2] = {0.0f, 0.0f};
    uint64_t step_ = 0;
    bool trainable_ = false;
};

}  // namespace sur
}  // namespace ion
