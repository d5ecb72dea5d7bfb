// Learned warm start: per-face features, training targets, the initial
// guess of both solves from the network, and the C ABI of the network.
//
// Features of face i (16): log c_salt, hs, u*, log u_z at the bottom and top
// cell, the bottom cell's sublimation coefficient, log area, log perimeter,
// then for the three edges ordered by the inflow they carry (E_j max(0,
// -u_hat . m_j)): inflow share of the perimeter and log c_salt of the upwind
// face (the face itself on the domain boundary), the outflow share, and T.
// Targets (11): log(c_z + 1e-9) of the 10 cells and asinh(1e4 qdep).
// The warm start writes x0 = max(exp(y_z) - 1e-9, 0) and qdep0 = 1e-4 sinh(y_10);
// both solves still iterate to ||b - A x|| <= tol ||b||.
#include <cmath>

#include "context.hpp"

namespace pb {
namespace {

constexpr double kFloor = 1e-9;
constexpr double kQscale = 1e4;
constexpr double kPi = 3.14159265358979323846;

__global__ void __launch_bounds__(kBlock) k_features(int64_t F, int L, const double* __restrict__ fscal,
                                                     const double* __restrict__ u_z, const double* __restrict__ csubl,
                                                     const double* __restrict__ area, const double* __restrict__ elen,
                                                     const double* __restrict__ nx, const double* __restrict__ ny,
                                                     const int32_t* __restrict__ neigh,
                                                     const double* __restrict__ vw_dir, const double* __restrict__ t,
                                                     float* __restrict__ X) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    double h = 450. - vw_dir[i];
    if (h > 360.) h -= 360.;
    const double vx = -cos(h * kPi / 180.), vy = -sin(h * kPi / 180.);
    const double cs = fscal[i];
    double per = 0, flux_in[3], lcs[3], out = 0;
    for (int j = 0; j < 3; ++j) {
        const double E = elen[j * F + i];
        const double um = vx * nx[j * F + i] + vy * ny[j * F + i];
        per += E;
        flux_in[j] = E * fmax(0.0, -um);
        out += E * fmax(0.0, um);
        const int32_t n = neigh[j * F + i];
        lcs[j] = log((n >= 0 ? fscal[n] : cs) + kFloor);
    }
    // order edges by inflow, largest first
    int o[3] = {0, 1, 2};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2 - a; ++b)
            if (flux_in[o[b]] < flux_in[o[b + 1]]) {
                const int tmp = o[b];
                o[b] = o[b + 1];
                o[b + 1] = tmp;
            }
    float* x = X + i * 16;
    x[0] = (float)log(cs + kFloor);
    x[1] = (float)fscal[F + i];
    x[2] = (float)fscal[2 * F + i];
    x[3] = (float)log(u_z[i]);
    x[4] = (float)log(u_z[(int64_t)(L - 1) * F + i]);
    x[5] = (float)(csubl[i] * 1e3);
    x[6] = (float)log(area[i]);
    x[7] = (float)log(per);
    for (int k = 0; k < 3; ++k) {
        x[8 + 2 * k] = (float)(flux_in[o[k]] / per);
        x[9 + 2 * k] = (float)lcs[o[k]];
    }
    x[14] = (float)(out / per);
    x[15] = (float)(t[i] / 10.0);
}

__global__ void __launch_bounds__(kBlock) k_targets(int64_t F, int L, const double* __restrict__ xs,
                                                    const double* __restrict__ xd, float* __restrict__ Y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    float* y = Y + i * 11;
    for (int z = 0; z < 10; ++z) y[z] = (float)log(fmax(xs[(int64_t)z * F + i], 0.0) + kFloor);
    y[10] = (float)asinh(xd[i] * kQscale);
}

__global__ void __launch_bounds__(kBlock) k_apply(int64_t F, int L, const float* __restrict__ Y, double* __restrict__ xs,
                                                  double* __restrict__ xd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    const float* y = Y + i * 11;
    if (xs)
        for (int z = 0; z < L; ++z) xs[(int64_t)z * F + i] = fmax(exp((double)y[z]) - kFloor, 0.0);
    if (xd) xd[i] = sinh((double)y[10]) / kQscale;
}

}  // namespace

void warm_features(pbsm3d_ctx* c, const double* vw_dir, const double* t) {
    hipLaunchKernelGGL(k_features, dim3(grid_for(c->F)), dim3(kBlock), 0, c->stream, c->F, c->L, c->fscal.p, c->u_z.p,
                       c->csubl.p, c->area.p, c->elen.p, c->nx.p, c->ny.p, c->neigh.p, vw_dir, t, c->feat.p);
}

void training_pairs_into(pbsm3d_ctx* c, const double* vw_dir, const double* t, float* X, float* Y) {
    hipLaunchKernelGGL(k_features, dim3(grid_for(c->F)), dim3(kBlock), 0, c->stream, c->F, c->L, c->fscal.p, c->u_z.p,
                       c->csubl.p, c->area.p, c->elen.p, c->nx.p, c->ny.p, c->neigh.p, vw_dir, t, X);
    hipLaunchKernelGGL(k_targets, dim3(grid_for(c->F)), dim3(kBlock), 0, c->stream, c->F, c->L, c->susp.x.p,
                       c->dep.x.p, Y);
    PB_HIP(hipGetLastError());
}

void warm_apply(pbsm3d_ctx* c, int which) {
    hipLaunchKernelGGL(k_apply, dim3(grid_for(c->F)), dim3(kBlock), 0, c->stream, c->F, c->L, c->pred.p,
                       which == 0 ? c->susp.x.p : nullptr, which == 1 ? c->dep.x.p : nullptr);
}

}  // namespace pb

// ---------------------------------------------------------------------------
// C ABI of the network
// ---------------------------------------------------------------------------
struct pbsm3d_mlp {
    pb::Mlp* m = nullptr;
    std::string err;
    ~pbsm3d_mlp() { delete m; }
};

namespace {
template <typename Fn>
int guard(pbsm3d_mlp* h, Fn&& fn) {
    try {
        fn();
        return PBSM3D_OK;
    } catch (const pb::Error& e) {
        if (h) h->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (h) h->err = e.what();
        return PBSM3D_EINVAL;
    }
}
}  // namespace

extern "C" {

int pbsm3d_mlp_create(pbsm3d_mlp** out, int device, uint64_t seed) {
    if (!out) return PBSM3D_EINVAL;
    *out = nullptr;
    auto* h = new pbsm3d_mlp();
    const int rc = guard(h, [&] { h->m = new pb::Mlp(device, seed); });
    if (rc != PBSM3D_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return PBSM3D_OK;
}

void pbsm3d_mlp_destroy(pbsm3d_mlp* h) { delete h; }

int64_t pbsm3d_mlp_param_count(void) { return pb::Mlp::param_count(); }

int pbsm3d_mlp_get_params(pbsm3d_mlp* h, float* host) {
    if (!h || !host) return PBSM3D_EINVAL;
    return guard(h, [&] { h->m->get_params(host, false); });
}

int pbsm3d_mlp_set_params(pbsm3d_mlp* h, const float* host) {
    if (!h || !host) return PBSM3D_EINVAL;
    return guard(h, [&] { h->m->set_params(host); });
}

int pbsm3d_mlp_set_normalization(pbsm3d_mlp* h, const float* host) {
    if (!h || !host) return PBSM3D_EINVAL;
    return guard(h, [&] { h->m->set_normalization(host); });
}

int pbsm3d_mlp_train(pbsm3d_mlp* h, const float* X, const float* Y, int64_t n, int32_t epochs, int32_t batch, float lr,
                     uint64_t seed, double* loss) {
    if (!h || !X || !Y || n <= 0 || epochs < 1 || batch < 1) return PBSM3D_EINVAL;
    return guard(h, [&] {
        const double l = h->m->train(X, Y, n, epochs, batch, lr, seed);
        if (loss) *loss = l;
    });
}

int pbsm3d_mlp_predict(pbsm3d_mlp* h, const float* X, float* Y, int64_t n) {
    if (!h || !X || !Y || n <= 0) return PBSM3D_EINVAL;
    return guard(h, [&] {
        h->m->predict(X, Y, n, h->m->stream);
        PB_HIP(hipStreamSynchronize(h->m->stream));
    });
}

int pbsm3d_mlp_predict_path(pbsm3d_mlp* h, const float* X, float* Y, int64_t n, int path) {
    if (!h || !X || !Y || n <= 0 || path < 0 || path > 2) return PBSM3D_EINVAL;
    if (path == 1 && !h->m->fused_inference_ok()) return PBSM3D_EINVAL;
    return guard(h, [&] {
        if (path == 0)
            h->m->predict(X, Y, n, h->m->stream);
        else if (path == 1)
            h->m->predict_fused(X, Y, n, h->m->stream);
        else
            h->m->predict_layers(X, Y, n, h->m->stream);
        PB_HIP(hipStreamSynchronize(h->m->stream));
    });
}

int pbsm3d_mlp_gradients(pbsm3d_mlp* h, const float* X, const float* Y, int64_t n, float* grads, double* loss) {
    if (!h || !X || !Y || n <= 0) return PBSM3D_EINVAL;
    return guard(h, [&] {
        const double l = h->m->gradients(X, Y, n, grads);
        if (loss) *loss = l;
    });
}

const char* pbsm3d_mlp_last_error(const pbsm3d_mlp* h) { return h ? h->err.c_str() : ""; }

int pbsm3d_training_pairs(pbsm3d_ctx* c, float* X, float* Y) {
    if (!c || !X || !Y || c->L != 10 || !c->last_vw_dir) return PBSM3D_EINVAL;
    try {
        PB_HIP(hipSetDevice(c->device));
        pb::training_pairs_into(c, c->last_vw_dir, c->last_t, X, Y);
        PB_HIP(hipStreamSynchronize(c->stream));
        return PBSM3D_OK;
    } catch (const pb::Error& e) {
        c->err = e.what();
        return e.code;
    }
}

int pbsm3d_set_warm_start(pbsm3d_ctx* c, pbsm3d_mlp* h) {
    if (!c) return PBSM3D_EINVAL;
    if (h && c->L != 10) {
        c->err = "the learned warm start predicts 10 layers: nLayer must be 10";
        return PBSM3D_EINVAL;
    }
    c->warm = h ? h->m : nullptr;
    return PBSM3D_OK;
}

}  // extern "C"
