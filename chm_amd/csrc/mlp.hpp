// Per-face network of the learned warm start (mlp.hip).
#pragma once

#include "kernels.hpp"

namespace pb {

struct MlpDims {
    static constexpr int IN = 16;   // features of a face and its three neighbours
    static constexpr int H = 128;   // hidden width (three hidden layers)
    static constexpr int OUT = 11;  // 10 layer concentrations + deposition flux
    static constexpr int OUTP = 16; // output padded to one MFMA tile
    static int din(int l);
    static int dout(int l);
};

class Mlp {
  public:
    Mlp(int device, uint64_t seed);
    ~Mlp();
    // X: [n x IN] f32, Y: [n x OUT] f32 device arrays (raw units; normalised inside)
    double train(const float* X, const float* Y, int64_t n, int epochs, int batch, float lr, uint64_t seed);
    // inference: the fused one-launch kernel (or layer by layer through the training
    // GEMM; the two agree bit for bit)
    void predict(const float* X, float* Y, int64_t n, hipStream_t s);
    void predict_fused(const float* X, float* Y, int64_t n, hipStream_t s);
    void predict_layers(const float* X, float* Y, int64_t n, hipStream_t s);
    bool fused_inference_ok();
    double gradients(const float* Xh, const float* Yh, int64_t n, float* grads_host);  // MSE gradients, host arrays
    static int64_t param_count();
    void get_params(float* host, bool grads);
    void set_params(const float* host);
    void set_normalization(const float* host);  // x_mean[IN], x_std[IN], y_mean[OUT], y_std[OUT]

    int device = 0;
    hipStream_t stream = nullptr;
    int fused_ok = -1;       // fused inference set up on this device (-1: not yet)
    int infer_blocks = 512;  // persistent grid of the fused inference (two workgroups per CU)

  private:
    void ensure(int64_t n);
    void prepare(const float* X, const float* Y, int64_t n, const int64_t* rows);
    // keep_z: also store the FP32 pre-activations of the hidden layers (for backward)
    void forward_prepared(int64_t n, bool keep_z = true);
    // loss_acc (device, may be null): the batch's mean squared error is added there on the
    // device and 0 is returned; null: the loss is returned (a host round trip)
    double backward_prepared(int64_t n, float* dY_f32, double* loss_acc = nullptr);
    void adam_step(float lr);
    const float* xmu() const;
    const float* xsd() const;
    const float* ymu() const;
    const float* ysd() const;

    DevBuf<float> W[4], bvec[4], gW[4], gb[4], mW[4], vW[4], mb[4], vb[4];
    DevBuf<uint16_t> Wb[4];
    DevBuf<uint8_t> keepW4, keepb4;
    DevBuf<float> norm;
    int64_t cap = 0;
    DevBuf<uint16_t> Xb, Ab[3], dZb[3], dYb;
    DevBuf<float> Z[3], Yp, Tn, partial;
    DevBuf<float> splitws;  // split-K partial products of the weight gradients
    DevBuf<float> colpart;  // per-workgroup column sums of the bias gradients
    DevBuf<double> loss_dev;  // the epoch's summed batch losses (training)
    DevBuf<int64_t> idx;
    int64_t t_adam = 0;
};

}  // namespace pb
