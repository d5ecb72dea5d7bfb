// The context behind a pbsm3d_ctx handle: device copies of the mesh, the two
// linear systems, per-step fields, solver workspaces and instrumentation.
#pragma once

#include <string>

#include "krylov.hpp"
#include "mlp.hpp"

namespace pb {
constexpr int kCoarseAdj = 16;        // tiles one coarse tile may touch, itself included
constexpr int kCoarseMaxTiles = 256;  // coarse unknowns (Gauss-Jordan inverse in one workgroup)
// The learned coarse space of the suspension solve: T tiles of faces (index ranges),
// the basis w (the network's predicted profile), the Galerkin matrix W^T A W and its
// inverse.
struct CoarseSpace {
    int T = 0;
    bool ready = false;  // G / Ginv hold the current step's coarse system
    DevBuf<int64_t> tile_off;  // [T + 1] face ranges
    DevBuf<int32_t> tile_of_face, tile_of_pos, adj;  // [F], [F], [T x kCoarseAdj]
    DevBuf<double> w, wpos, rref;  // [N] basis (reference / position order), residual in reference order
    DevBuf<double> G, M, Ginv, c, y;
};
}  // namespace pb

struct pbsm3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    int64_t F = 0;
    int L = 0;
    pbsm3d_config cfg{};
    double dz = 0;
    bool enable_veg = false;

    pb::DevBuf<double> nx, ny, elen, area, cz, canopy, lai, stalk_n, stalk_d, dep_w;
    pb::DevBuf<int32_t> neigh;
    pb::DevBuf<uint16_t> slots;
    pb::DevBuf<uint8_t> water;
    pb::System susp, dep;
    pb::DevBuf<double> u_z, csubl, Qsalt, Qsusp, Qsubl, Qsubl_mass, sum_subl, drift_mass, sum_drift, more;
    pb::DevBuf<double> blowprob;  // blowingsnow_probability (use_PomLi_probability)
    pb::DevBuf<double> z0_face, settle_face;  // per-face z0_snow / settling_scale of a batch union (else empty)
    pb::DevBuf<double> fscal;  // [3F] per-face c_salt, hs, u* of the last step
    pb::DevBuf<uint8_t> salt_flag;
    pb::DevBuf<double> forcing_stage;
    pb::ColumnFactors colf;
    pb::ColorOrder co;
    pb::DiluFactors<double> dl_s, dl_d;
    pb::DiluFactors<float> dl32;
    pb::DevBuf<double> invdiag_s, invdiag_d, r_s, r_d;
    pb::DevBuf<double> bp, xp;  // suspension b and x in the D-ILU position order
    pb::DevBuf<double> split64[3];  // work vectors of the split BiCGStab (FP64 / MIXED)
    pb::DevBuf<float> split32[3];
    pb::DevBuf<double> dsplit[5];  // deposition b, x (position order) and split BiCGStab work vectors
    pb::BicgWork<double> bwd64;    // deposition BiCGStab workspace
    // position-order SpMV launch shape per precision (index 0: FP64, 1: FP32), -1 until tuned
    int spmv_variant[2] = {-1, -1};
    float spmv_tune_ms[2][pb::kPosSpmvVariants] = {};
    // D-ILU sweeps with the XCD workgroup mapping (1) or without (0), -1 until tuned
    // D-ILU sweep mode per precision: bit 0 the XCD workgroup mapping, bit 1 (split
    // operator only) the non-temporal factor loads; -1 until tuned
    int dilu_mode[2] = {-1, -1};
    // shadow-residual restart threshold of the FP32 BiCGStab runs (bicg_stage; 0: never)
    double restart_eps32 = pb::kBicgRestartEps32;
    float dilu_tune_ms[2][pb::kSplitModes] = {};
    pb::DevBuf<float> val32, r32, d32;
    pb::GmresWork<double> ws64, wd64;
    pb::GmresWork<float> ws32;
    pb::BicgWork<double> bw64;
    pb::BicgWork<float> bw32;
    pb::DevBuf<double> red_partial, red_out;
    pb::HostBuf<double> red_host;
    // learned warm start
    pb::Mlp* warm = nullptr;          // not owned
    pb::DevBuf<float> feat, pred;     // [F x 16], [F x 11]
    const double* last_vw_dir = nullptr;  // this step's vw_dir and t, context-owned copies (device)
    const double* last_t = nullptr;
    pb::CoarseSpace coarse;  // learned coarse correction (cfg.coarse_tiles > 0)
    pb::Profiler prof;
    std::string err;

    ~pbsm3d_ctx() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace pb {
// the argument block of k_assemble for this context and step (fdev: the nine forcing
// fields on the device, NULL where absent)
AssembleArgs assemble_args(pbsm3d_ctx* c, const double* const fdev[9], int32_t dt_s);
// features of the current step into c->feat (needs the assembled step)
void warm_features(pbsm3d_ctx* c, const double* vw_dir, const double* t);
// training pairs of the last step over all c->F faces: X [F x 16] features (vw_dir, t the
// step's forcing on the device), Y [F x 11] targets from the solved x of both systems
void training_pairs_into(pbsm3d_ctx* c, const double* vw_dir, const double* t, float* X, float* Y);
// x0 of the suspension (which = 0) or deposition (which = 1) solve from c->pred
void warm_apply(pbsm3d_ctx* c, int which);
// learned coarse correction: tiles and adjacency (at create, from the host mesh), the step's
// basis and Galerkin inverse (after the learned start wrote x0), and one correction
// x_pos += W (W^T A W)^-1 W^T r (r_pos the FP64 residual, both in position order)
void coarse_setup(pbsm3d_ctx* c, int T, const int32_t* neigh_face_major);
void coarse_prepare(pbsm3d_ctx* c);
void coarse_correct(pbsm3d_ctx* c, hipStream_t s, const double* r_pos, double* x_pos);
}  // namespace pb
