/*
 * pbsm3d.h -- C ABI of the MI355X PBSM3D blowing-snow step.
 *
 * One context = one mesh on one GPU, one HIP stream, all device memory owned by
 * the context.  A context is not thread-safe; separate contexts are independent.
 *
 * The entry points replace, for the C++ PBSM3D module, the Trilinos-backed pieces
 * of the reference step:
 *   pbsm3d_create          PBSM3D::init            src/modules/PBSM3D.cpp:221-398
 *                          + NearestNeighborProblem ctor  src/math/LinearAlgebra.cpp:31-197
 *   pbsm3d_step            PBSM3D::run             src/modules/PBSM3D.cpp:400-1748
 *                          (zeroSystem/sumInto/Solve  src/math/LinearAlgebra.cpp:199-252)
 *   pbsm3d_export_system   NearestNeighborProblem::writeSystemMatrixMarket
 *                                                  src/math/LinearAlgebra.cpp:277-287
 *   pbsm3d_get_state /     PBSM3D::checkpoint / load_checkpoint
 *   pbsm3d_set_state                               src/modules/PBSM3D.cpp:1753-1773
 *   pbsm3d_destroy         PBSM3D::~PBSM3D         src/modules/PBSM3D.cpp:1750-1751
 *
 * Conventions: 0 = success, negative = error (pbsm3d_last_error gives text).
 * Per-face arrays have length F; per-cell arrays have length nLayer*F in the
 * reference's layer-major order (row r = z*F + face, LinearAlgebra.cpp:43-63).
 * Per-edge arrays have length 3*F, face-major: [3*face + j] for edge j.
 * All floating point is IEEE double unless a name says otherwise.
 */
#ifndef PBSM3D_H
#define PBSM3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBSM3D_ABI_VERSION 7

/* return codes */
#define PBSM3D_OK 0
#define PBSM3D_EINVAL (-1)   /* invalid argument / config (reference: module_error on bad config) */
#define PBSM3D_EHIP (-2)     /* HIP runtime error */
#define PBSM3D_ENOCONV (-3)  /* Krylov solve did not converge (reference: module_error, PBSM3D.cpp:1447,1704) */
#define PBSM3D_ENAN (-4)     /* non-finite value detected */
#define PBSM3D_ENOMEM (-5)   /* device allocation failed */

/* arithmetic of the linear solves */
#define PBSM3D_FP64 0   /* FP64 SpMV, preconditioner and Krylov */
#define PBSM3D_MIXED 1  /* FP32 SpMV/preconditioner inner solves, FP64 refinement, FP64 dots/residuals */

/* preconditioner of the suspension solve */
#define PBSM3D_PC_NONE 0
#define PBSM3D_PC_JACOBI 1
#define PBSM3D_PC_COLUMN 2   /* per-face vertical tridiagonal block (Thomas, nLayer unknowns) */
#define PBSM3D_PC_ILU0 3     /* ILU(0) on the CSR in the natural order (CPU oracle only; the GPU
                                path uses the multicolour form PBSM3D_PC_MCILU0) */
#define PBSM3D_PC_ILUT 4     /* ILUT(1e-4, fill 3) as the reference's Ifpack2 setup (CPU oracle only) */
#define PBSM3D_PC_MCILU0 5   /* D-ILU(0) in the multicolour order (face colour, layer, face): faces are
                                greedily coloured so that no two neighbours share a colour; each colour is
                                one parallel phase in which a thread factors / solves one face column */

/* initial guess of the suspension solve */
#define PBSM3D_WS_ZERO 0     /* x0 = 0, as zeroSystem (LinearAlgebra.cpp:199-206) */
#define PBSM3D_WS_LEARNED 1  /* per-face network prediction (see pbsm3d_mlp_*) */

typedef struct pbsm3d_mesh {
    int64_t n_faces;
    const int32_t* neigh;       /* [3F] face::neighbor(j) index, -1 = none          */
    const double* normal_x;     /* [3F] face::edge_unit_normal(j).x                 */
    const double* normal_y;     /* [3F] face::edge_unit_normal(j).y                 */
    const double* edge_length;  /* [3F] face::edge_length(j)                        */
    const double* area;         /* [F]  face::get_area()                            */
    const double* center_x;     /* [F]  face::center().x                            */
    const double* center_y;     /* [F]  face::center().y                            */
    const double* center_z;     /* [F]  face::get_z()                               */
    const double* canopy_height;  /* [F] veg_attribute("CanopyHeight") or NULL     */
    const double* lai;            /* [F] veg_attribute("LAI") or NULL              */
    const double* stalk_number;   /* [F] veg_attribute("stalk_number") or NULL     */
    const double* stalk_diameter; /* [F] veg_attribute("stalk_diameter") or NULL;
                                     either NULL: N = 1, dv = 0.8 on every face
                                     (PBSM3D.cpp:302-309)                          */
    const uint8_t* is_water;      /* [F] module_base::is_water(face) or NULL       */
} pbsm3d_mesh;

/* config.PBSM3D keys (PBSM3D.cpp:123-145, 223-258) plus solver settings that the
 * reference hard-codes (LinearAlgebra.cpp:165-185) and extension keys whose
 * defaults reproduce the reference. */
typedef struct pbsm3d_config {
    int32_t nLayer;               /* 10  (>= 2)                                    */
    int32_t do_fixed_settling;    /* 0                                             */
    double settling_velocity;     /* 0.5 m/s, must be >= 0                         */
    int32_t do_sublimation;       /* 1                                             */
    int32_t do_lateral_diff;      /* 1                                             */
    double smooth_coeff;          /* 820                                           */
    double min_sd_trans;          /* 0.1 m                                         */
    double cutoff;                /* 0.3 m                                         */
    double snow_diffusion_const;  /* 0.3                                           */
    int32_t enable_veg;           /* 1 (turned off when the mesh has no veg data)  */
    int32_t use_tanh_fetch;       /* 1                                             */
    int32_t use_exp_fetch;        /* 0                                             */
    int32_t use_R94_lambda;       /* 1                                             */
    /* extension keys */
    double z0_snow;               /* 0.01 = Snow::Z0_SNOW                          */
    double settling_scale;        /* 1.0: multiplies the settling velocity         */
    /* solver */
    double tol;                   /* 1e-8: ||b - A x||_2 / ||b||_2                 */
    int32_t max_iters;            /* 1000                                          */
    int32_t restart;              /* 30                                            */
    int32_t precision;            /* PBSM3D_FP64                                   */
    int32_t precond;              /* PBSM3D_PC_MCILU0                              */
    int32_t warm_start;           /* PBSM3D_WS_ZERO                                */
    int32_t dep_precond;          /* PBSM3D_PC_MCILU0 (deposition system)          */
    int32_t ortho_passes;         /* 1 (default): one classical Gram-Schmidt pass;
                                     2: a second pass when the norm drops below
                                     1/sqrt(2) (DGKS re-orthogonalisation)          */
    int32_t block_size;           /* reserved (256)                                */
    int32_t krylov;               /* PBSM3D_KRYLOV_GMRES (reference) or _BICGSTAB  */
    double inner_rtol;            /* MIXED: relative residual each FP32 inner solve
                                     aims for; 0 = default (split BiCGStab: 2e-3
                                     from x0 = 0, 2e-4 from the learned start;
                                     1e-4 otherwise)                               */
    /* reference keys that are off by default */
    int32_t use_PomLi_probability;  /* 0 (PBSM3D.cpp:124, 848-865): c_salt scaled by the
                                       probability of blowing snow of Essery, Li & Pomeroy
                                       1999 eqs 10-14 from the forcing's p_snow_hours; the
                                       probability is the blowingsnow_probability output   */
    int32_t z0_ustar_coupling;      /* 0 (PBSM3D.cpp:125, 691-737): u* of saltation
                                       candidates from the Li & Pomeroy 2000 eq. 5 fixed
                                       point u* = u2 k / ln(2 / (c u*^2 + lambda/2)), and
                                       z0 = c u*^2 + lambda/2 while saltating (a bracketing
                                       root search per face)                              */
    int32_t rouault_diffusion_coef; /* 0 (PBSM3D.cpp:258, 1164-1172): vertical diffusivity
                                       coefficient 1 / (1 + w^2 / (1.56 u*^2)) per cell in
                                       place of snow_diffusion_const                        */
    int32_t coarse_tiles;           /* 0 (extension): with the learned warm start and the
                                       split BiCGStab, T in [1, 256] adds the learned coarse
                                       correction: at the start of every refinement cycle
                                       x += W (W^T A W)^-1 W^T (b - A x), W = the network's
                                       predicted profile on T tiles (face-index ranges)     */
} pbsm3d_config;

/* Krylov method of the suspension solve (the deposition solve is GMRES(restart)) */
#define PBSM3D_KRYLOV_GMRES 0     /* restarted GMRES(restart), as Belos "GMRES"     */
#define PBSM3D_KRYLOV_BICGSTAB 1  /* BiCGStab: short recurrences, 2 SpMV/iteration  */
#define PBSM3D_KRYLOV_BICGSTAB_SPLIT 2  /* BiCGStab on the split D-ILU system (Eisenstat's
                                          trick: each operator application is one backward
                                          and one forward sweep, no SpMV) inside FP64
                                          refinement; precond MCILU0, nLayer <= 16.  With
                                          dep_precond MCILU0 the deposition system is solved
                                          the same way (FP64) */

typedef struct pbsm3d_forcing {
    const double* U_R;            /* [F] m/s at Z_U_R = 50 m                       */
    const double* U_2m_above_srf; /* [F] m/s                                       */
    const double* vw_dir;         /* [F] degrees (bearing the wind blows from)     */
    const double* swe;            /* [F] kg m-2 (-9999/NaN -> 0)                   */
    const double* t;              /* [F] deg C                                     */
    const double* rh;             /* [F] %                                         */
    const double* snowdepthavg;   /* [F] m (-9999/NaN -> 0)                        */
    const double* fetch;          /* [F] m, or NULL when no fetch mode is on       */
    const double* p_snow_hours;   /* [F] hours since snowfall, read by
                                     use_PomLi_probability only (else may be NULL)  */
    int32_t on_device;            /* 1: all pointers are device (HBM) pointers     */
} pbsm3d_forcing;

typedef struct pbsm3d_out {
    /* per-face module outputs, NULL to skip (the context keeps its own copy)      */
    double* Qsalt;
    double* Qsusp;
    double* Qsubl;
    double* Qsubl_mass;
    double* sum_subl;
    double* drift_mass;
    double* sum_drift;
    double* pbsm_more_than_avail;
    double* blowingsnow_probability; /* written on saltating faces with
                                        use_PomLi_probability, else kept (-9999 at first) */
    /* solver fields for parity checks, NULL to skip                               */
    double* conc;   /* [nLayer*F] suspension solution (after no clamping)          */
    double* qdep;   /* [F] deposition solution                                     */
    int32_t on_device;
} pbsm3d_out;

typedef struct pbsm3d_stats {
    int32_t suspension_present;
    int32_t deposition_present;
    int32_t susp_iters;           /* total Krylov iterations (all restarts/cycles)  */
    int32_t dep_iters;
    double susp_resid;            /* final ||b - A x|| / ||b|| (explicit)           */
    double dep_resid;
    double susp_rhs_max;          /* ||b||_inf (getRhsMax)                          */
    double dep_rhs_max;
    int32_t susp_refinements;     /* outer FP64 refinement steps (MIXED)            */
    int32_t susp_restarts;        /* shadow-residual restarts of the FP32 BiCGStab  */
    double ms_assembly;           /* HIP-event times of the last step, ms           */
    double ms_susp_solve;
    double ms_post;
    double ms_dep_solve;
    double ms_total;
    double susp_resid0;           /* ||b - A x0|| / ||b|| of the suspension start   */
    int32_t susp_cold_fallback;   /* 1: the warm start was worse than x0 = 0 and    */
    int32_t reserved2;            /*    the solve restarted from zero               */
    double susp_resid_coarse;     /* ||b - A x|| / ||b|| after the first learned coarse
                                     correction (0: not applied)                     */
} pbsm3d_stats;

typedef struct pbsm3d_ctx pbsm3d_ctx;

void pbsm3d_config_default(pbsm3d_config* cfg);

/* Face geometry from vertices, restating the face API the module reads in init():
 * edge_unit_normal, edge_length, get_area (signed 2-D area, or the "area"
 * parameter when given), center (triangulation.hpp:1443-1498, 1577-1589,
 * 1774-1780, 1830-1856).  Host arrays; outputs sized as in pbsm3d_mesh. */
int pbsm3d_face_geometry(int64_t n_faces, const double* vertex_xyz, const int32_t* elem,
                         const double* area_param, double* normal_x, double* normal_y,
                         double* edge_length, double* area, double* center_x, double* center_y,
                         double* center_z);

/* init(): copies geometry to HBM, builds both CSR patterns (bit-exact with
 * LinearAlgebra.cpp:67-149), allocates solver workspaces on device `device`. */
int pbsm3d_create(pbsm3d_ctx** out, const pbsm3d_mesh* mesh, const pbsm3d_config* cfg, int device);

/* run(): one full timestep of dt_s seconds; outputs synchronised on return. */
int pbsm3d_step(pbsm3d_ctx* ctx, const pbsm3d_forcing* forcing, int32_t dt_s, pbsm3d_out* out,
                pbsm3d_stats* stats);

/* nnz of the suspension (which=0) or deposition (which=1) system. */
int64_t pbsm3d_system_nnz(const pbsm3d_ctx* ctx, int which);

/* Copy the last assembled system (pattern + values + RHS) to host buffers;
 * any pointer may be NULL.  row_ptr has n_rows+1 entries. */
int pbsm3d_export_system(pbsm3d_ctx* ctx, int which, int32_t* row_ptr, int32_t* col, double* val,
                         double* rhs);

/* Checkpoint state: sum_drift (checkpointed by the reference) and sum_subl
 * (module state the reference does not checkpoint), host arrays of F. */
int pbsm3d_get_state(pbsm3d_ctx* ctx, double* sum_drift, double* sum_subl);
int pbsm3d_set_state(pbsm3d_ctx* ctx, const double* sum_drift, const double* sum_subl);

void pbsm3d_destroy(pbsm3d_ctx* ctx);
const char* pbsm3d_last_error(const pbsm3d_ctx* ctx); /* ctx may be NULL: last global error */
int pbsm3d_abi_version(void);

/* Instrumentation (no counterpart in the reference, whose only timing is the
 * per-timestep timer of core.cpp:2411-2437): with profiling on, each launch of
 * the kernel classes below is bracketed by HIP events recorded on the
 * context's own stream, and pbsm3d_profile_read returns the launch count and
 * the summed device time since the last reset. */
#define PBSM3D_K_ASSEMBLE 0    /* coefficients + suspension CSR assembly       */
#define PBSM3D_K_SPMV64 1      /* FP64 suspension SpMV (Arnoldi)               */
#define PBSM3D_K_SPMV32 2      /* FP32 suspension SpMV (mixed inner cycles)    */
#define PBSM3D_K_PRECOND 3     /* suspension preconditioner apply              */
#define PBSM3D_K_ORTHO 4       /* batched Gram-Schmidt dots + update           */
#define PBSM3D_K_RESIDUAL 5    /* FP64 residual b - A x (fused norm)           */
#define PBSM3D_K_DEPOSITION 6  /* post-processing + deposition assembly + solve */
#define PBSM3D_K_WARMSTART 7   /* learned warm start (network forward)         */
#define PBSM3D_K_FACTOR 8      /* suspension D-ILU factorisation (once a step) */
#define PBSM3D_K_COUNT 9
int pbsm3d_profile_enable(pbsm3d_ctx* ctx, int enable);
int pbsm3d_profile_read(pbsm3d_ctx* ctx, int kernel, int64_t* launches, double* ms_total, int reset);
/* Restricts the bracketing to the classes whose bit (1 << PBSM3D_K_*) is set
 * (default: all).  Each bracket adds two event records to the stream, a few
 * microseconds of device idle time, so a timing run brackets only the classes
 * it reports. */
int pbsm3d_profile_classes(pbsm3d_ctx* ctx, uint32_t mask);
/* Brackets only every `every`-th launch of each class (default 1: all of them); the counts
 * and times read back are those of the bracketed launches, so launches are homogeneous
 * samples of their class while the idle time of the event records drops by that factor. */
int pbsm3d_profile_sample(pbsm3d_ctx* ctx, int every);

/* Launch shape of the suspension SpMV (no counterpart in the reference).  With
 * the multicolour D-ILU preconditioner the suspension solve runs in the
 * preconditioner's position order and its SpMV has several launch shapes (one
 * thread per row or per face column, several block sizes, with or without a
 * workgroup order that keeps neighbouring faces on one XCD); the first solve of
 * each precision times them all with HIP events and keeps the fastest.
 * which: 0 = FP64 operator, 1 = FP32 operator (MIXED).  *chosen = -1 until
 * tuned; ms[v] (n entries, may be NULL) = average time of shape v in ms (0 if
 * not tried). */
int pbsm3d_spmv_tuning(const pbsm3d_ctx* ctx, int which, int32_t* chosen, float* ms, int32_t n);
/* Fixes the launch shape (0 .. PBSM3D_SPMV_VARIANTS-1) instead of the tuner's
 * choice, or -1 to tune again at the next solve. */
#define PBSM3D_SPMV_VARIANTS 10
int pbsm3d_spmv_select(pbsm3d_ctx* ctx, int which, int variant);
/* The same for the multicolour D-ILU sweeps.  *mode (-1 until tuned) is the
 * sweep mode the first solve found fastest: bit 0 = the workgroup order that
 * keeps neighbouring faces on one XCD; with the split operator
 * (PBSM3D_KRYLOV_BICGSTAB_SPLIT) also bit 1 = the factor arrays read with the
 * non-temporal hint and bit 2 = layer-parallel sweeps (a workgroup per 64 face
 * columns, one wave per layer, the vertical recurrence from LDS) instead of one
 * thread per column; the other solvers tune bit 0 alone.
 * ms[PBSM3D_PRECOND_MODES] = average apply time of each mode (ms; 0 if not
 * tried).  pbsm3d_precond_select fixes the mode (0 .. 1, or 0 .. 7 with the
 * split operator) or re-tunes (-1). */
#define PBSM3D_PRECOND_MODES 8
int pbsm3d_precond_tuning(const pbsm3d_ctx* ctx, int which, int32_t* mode, float* ms);
int pbsm3d_precond_select(pbsm3d_ctx* ctx, int which, int mode);

/* Restart threshold of the shadow residual in the FP32 BiCGStab runs (MIXED with
 * PBSM3D_KRYLOV_BICGSTAB or _BICGSTAB_SPLIT; no counterpart in the reference, whose
 * GMRES has no shadow vector): when |rh.r| < eps ||rh|| ||r||, a near breakdown in
 * which rho keeps few significant digits, the run continues with rh = p = r.  Default
 * 1e-5; 0 disables it; eps must lie in [0, 1).  stats.susp_restarts counts them. */
int pbsm3d_bicg_restart(pbsm3d_ctx* ctx, double eps);

/* Face colours of the multicolour D-ILU preconditioner (host computation, no
 * GPU needed; no counterpart in the reference): neighbouring faces never share
 * a colour.  neigh as in pbsm3d_mesh; color[F] out.  Returns the number of
 * colours (3 on triangulations unless a Kempe swap is blocked), < 0 on error. */
int pbsm3d_face_colors(int64_t n_faces, const int32_t* neigh, int32_t* color);

/* Learned warm start (an extension; the reference always starts from x0 = 0,
 * LinearAlgebra.cpp:199-206).  A per-face network 16 -> 128 -> 128 -> 128 -> 11
 * (ReLU, BF16 operands, FP32 accumulation on the matrix cores) maps features of
 * a face and its three neighbours to its 10-layer concentration profile and
 * deposition flux; with cfg.warm_start = PBSM3D_WS_LEARNED and a network
 * attached, both solves start from the prediction and still iterate to
 * ||b - A x|| <= tol ||b||.  The network is trained on the GPU by its own
 * forward/backward kernels on (features, solution) pairs of FP64 solves.
 * Parameters are flat f32: W1[128x16], b1, W2[128x128], b2, W3[128x128], b3,
 * W4[16x128] (rows >= 11 zero), b4[16].  Device arrays unless stated. */
typedef struct pbsm3d_mlp pbsm3d_mlp;
int pbsm3d_mlp_create(pbsm3d_mlp** out, int device, uint64_t seed);
void pbsm3d_mlp_destroy(pbsm3d_mlp* net);
int64_t pbsm3d_mlp_param_count(void);
int pbsm3d_mlp_get_params(pbsm3d_mlp* net, float* host_params);
int pbsm3d_mlp_set_params(pbsm3d_mlp* net, const float* host_params);
/* host: x_mean[16], x_std[16], y_mean[11], y_std[11] */
int pbsm3d_mlp_set_normalization(pbsm3d_mlp* net, const float* host_norm);
/* X [n x 16], Y [n x 11]; mini-batch Adam with MSE on normalised targets */
int pbsm3d_mlp_train(pbsm3d_mlp* net, const float* X, const float* Y, int64_t n, int32_t epochs, int32_t batch,
                     float lr, uint64_t seed, double* final_loss);
int pbsm3d_mlp_predict(pbsm3d_mlp* net, const float* X, float* Y, int64_t n);
/* The same through a chosen kernel path: 0 = as pbsm3d_mlp_predict (the fused
 * path), 1 = the fused one-launch kernel (all four layers per 64-sample tile,
 * activations in LDS), 2 = layer by layer through the training GEMM.  Paths 1
 * and 2 give the same outputs bit for bit. */
int pbsm3d_mlp_predict_path(pbsm3d_mlp* net, const float* X, float* Y, int64_t n, int path);
/* host arrays: MSE loss and its gradient w.r.t. all parameters on one batch */
int pbsm3d_mlp_gradients(pbsm3d_mlp* net, const float* X, const float* Y, int64_t n, float* host_grads, double* loss);
const char* pbsm3d_mlp_last_error(const pbsm3d_mlp* net);
/* features [F x 16] and solved targets [F x 11] of the context's last step (nLayer 10) */
int pbsm3d_training_pairs(pbsm3d_ctx* ctx, float* X, float* Y);
/* attach (or detach with NULL) the network used when cfg.warm_start = PBSM3D_WS_LEARNED */
int pbsm3d_set_warm_start(pbsm3d_ctx* ctx, pbsm3d_mlp* net);
/* The learned coarse space of the last step that used it (cfg.coarse_tiles = T > 0), for
 * parity checks (host arrays, each may be NULL): tile_of_face[F], the basis w[nLayer*F]
 * (the network's predicted profile, reference order) and the Galerkin matrix
 * G[T*T] = W^T A W (row-major, a tile with w = 0 has G[a][a] = 1). */
int pbsm3d_coarse_system(pbsm3d_ctx* ctx, int32_t* tile_of_face, double* w, double* G);

#ifdef __cplusplus
}
#endif

#endif /* PBSM3D_H */
