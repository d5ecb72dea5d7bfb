/* Batched parameter ensemble of the PBSM3D step on one MI355X.
 *
 * No single counterpart in the reference, which runs one PBSM3D module per model
 * instance (src/modules/PBSM3D.cpp:400-1748); an ensemble there is one CHM process per
 * member.  Here the members -- meshes of uneven size, each with its own PBSM3D
 * parameters -- advance together: members of equal nLayer are held as one union whose
 * matrices are block diagonal, each colour's sweep of the split D-ILU operator is one
 * launch for every member of the union, and the BiCGStab scalars (rho, alpha, omega,
 * the residual norm and the stop flag) are kept per member on the device, so each
 * member converges to its own ||b - A x|| <= tol ||b|| and then drops out of the work.
 * Every member's step is the reference's run(mesh&) on that member.
 *
 * Per-member parameters: nLayer, z0_snow, settling_scale.  Every other config field
 * must be the same for all members; the solver must be the split D-ILU BiCGStab
 * (krylov PBSM3D_KRYLOV_BICGSTAB_SPLIT, precond and dep_precond PBSM3D_PC_MCILU0), in
 * FP64 or MIXED precision.  The learned warm start (pbsm3d_batch_set_warm_start)
 * applies to the members with nLayer 10.
 */
#ifndef PBSM3D_BATCH_H
#define PBSM3D_BATCH_H

#include <stdint.h>

#include "pbsm3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pbsm3d_batch pbsm3d_batch;

/* meshes[m], cfgs[m] for m < n_members, as pbsm3d_create takes them (host arrays). */
int pbsm3d_batch_create(pbsm3d_batch** out, int32_t n_members, const pbsm3d_mesh* meshes,
                        const pbsm3d_config* cfgs, int device);

/* One timestep of every member: forcings[m] and outs[m] as pbsm3d_step takes them
 * (each forcing / output struct may hold host or device pointers); stats[m] (may be
 * NULL) per member: iterations, residuals and presence flags of that member, the ms_*
 * times of the whole batch step.  A member whose solve does not converge makes the
 * call return PBSM3D_ENOCONV (pbsm3d_batch_last_error names it). */
int pbsm3d_batch_step(pbsm3d_batch* b, const pbsm3d_forcing* forcings, int32_t dt_s, pbsm3d_out* outs,
                      pbsm3d_stats* stats);

/* attach (NULL: detach) the network of the learned warm start (nLayer-10 members) */
int pbsm3d_batch_set_warm_start(pbsm3d_batch* b, pbsm3d_mlp* net);

/* enable = 1 (default): the BiCGStab iterations after the first of every inner run are
 * replayed from two captured hipGraphs of two iterations each (captured at the first run
 * of each union and system); 0: every kernel launched on the stream.  The launches and
 * the results are the same either way. */
int pbsm3d_batch_use_graphs(pbsm3d_batch* b, int enable);

/* number of unions (distinct nLayer values) and launches of the last step's
 * suspension solves: *launches counts every kernel launch of the batched BiCGStab
 * iterations (all unions), *iterations the lock-step iterations they ran */
int pbsm3d_batch_info(const pbsm3d_batch* b, int32_t* unions, int64_t* launches, int64_t* iterations);

/* per-member checkpoint state, host arrays of that member's face count */
int pbsm3d_batch_get_state(pbsm3d_batch* b, int32_t member, double* sum_drift, double* sum_subl);
int pbsm3d_batch_set_state(pbsm3d_batch* b, int32_t member, const double* sum_drift, const double* sum_subl);

/* Training pairs of the learned warm start from member `member`'s last step (nLayer 10),
 * as pbsm3d_training_pairs gives them for one context: X [F_m x 16] features and Y
 * [F_m x 11] targets, device arrays of that member's face count.  Many forcing draws of
 * one mesh solved as the members of one batch give the network's training set in few
 * launches. */
int pbsm3d_batch_training_pairs(pbsm3d_batch* b, int32_t member, float* X, float* Y);

void pbsm3d_batch_destroy(pbsm3d_batch* b);
/* b NULL: the last pbsm3d_batch_create error */
const char* pbsm3d_batch_last_error(const pbsm3d_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* PBSM3D_BATCH_H */
