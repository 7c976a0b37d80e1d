// The learned FLD preconditioner: a 3-level U-Net (16/32/64 channels, 3x3 convolutions) on BF16 activations with
// FP32 accumulation on the matrix cores, its backward pass and Adam, all on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace sr {

typedef __bf16 bf16;

struct NetLayer {
    int cin, cout;   // real channels
    int cinp;        // input channels as stored (the first layer's 3 features padded to 8)
    int kp, kpb;     // forward K = 9 cinp and backward-data K = 9 cout, each rounded up to 32
    size_t w_off, b_off;
    bf16 *wt = nullptr;  // forward weights (cout, kp), k = tap * cinp + ci
    bf16 *wb = nullptr;  // backward-data weights (cin, kpb), k = tap' * cout + co, taps mirrored
};

struct NetStats {
    double loss = 0.0;
};

class Net {
  public:
    ~Net();
    int init(int max_w, int max_h);
    size_t param_count() const { return nparams; }
    int set_params(const float *host, size_t n);
    int get_params(float *host, size_t n);
    int get_grads(float *host, size_t n);
    // z = zadd + D^-1/2 s N(features(r)) (zadd may be null) on a w x h system (w, h multiples of 4)
    int apply(int w, int h, const float *own, const float *kx, const float *ky, const float *r, const float *zadd,
              float *z, hipStream_t s);
    // the energy-norm loss of the network's correction against e = A^-1 r and every parameter's gradient
    int loss_grad(int w, int h, const float *own, const float *kx, const float *ky, const float *r, const float *e,
                  double *loss, hipStream_t s);
    int adam(double lr, hipStream_t s);
    // an activation of the last forward pass as f32 on the host
    int activation(const char *name, float *host_out, size_t count);
    double bytes_per_apply(int w, int h) const;
    double flops_per_apply(int w, int h) const;

  private:
    int forward(int w, int h, const float *own, const float *kx, const float *ky, const float *r, hipStream_t s);
    int pack(hipStream_t s);
    int max_w = 0, max_h = 0;
    std::vector<NetLayer> L;
    size_t nparams = 0, out_off = 0;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    int adam_t = 0;
    bool packed = false;
    // activations, channels-last BF16, sized for max_w x max_h
    bf16 *x = nullptr, *e1a = nullptr, *e1 = nullptr, *p1 = nullptr, *e2a = nullptr, *e2 = nullptr, *p2 = nullptr,
         *e3a = nullptr, *e3 = nullptr, *d2a = nullptr, *d2 = nullptr, *d1a = nullptr, *d1 = nullptr;
    // gradients of the activations
    bf16 *g_d1 = nullptr, *g_d1a = nullptr, *g_cat1 = nullptr, *g_d2 = nullptr, *g_d2a = nullptr,
         *g_cat2 = nullptr, *g_e3 = nullptr, *g_e3a = nullptr, *g_p2 = nullptr, *g_e2 = nullptr, *g_e2a = nullptr,
         *g_p1 = nullptr, *g_e1 = nullptr, *g_e1a = nullptr;
    float *dis = nullptr, *err = nullptr, *aerr = nullptr, *gout = nullptr, *wpart = nullptr;
    double *part = nullptr, *lpart = nullptr, *opart = nullptr, *host = nullptr;
    float *scal = nullptr;   // s (the residual's scale) and the mean log coupling, on the device
    int nb_feat = 0, last_w = 0, last_h = 0;
};

} // namespace sr
