// Kernel argument structs and kernel declarations of the PBSM3D library.
#pragma once

#include "pbsm3d_internal.hpp"

namespace pb {

// module_base::is_nan (src/modules/module_base.hpp:471-479): a face variable is
// missing when it lies within 1e-5 of the -9999 default or is NaN.
__device__ __forceinline__ bool chm_missing(double v) { return fabs(v - -9999.0) < 1e-5 || isnan(v); }

// wavefront-wide sum (64 lanes)
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block-wide sum in a fixed order; result valid in thread 0; lds holds blockDim.x / 64 doubles
__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += lds[k];
    __syncthreads();
    return s;
}

// Workgroups of a launch are dispatched to the 8 XCDs round-robin (workgroup b
// runs on XCD b % 8), each XCD with its own L2.  With XCD = true a launch of
// launch_blocks(nb) workgroups covers nb logical blocks so that XCD j works on
// the j-th contiguous eighth of them: neighbouring faces, whose gathers share
// cache lines, then meet in one L2.  Logical blocks >= nb do nothing (the
// kernels' own range checks drop them).
constexpr int kXcd = 8;
template <bool XCD>
inline int64_t launch_blocks(int64_t nb) {
    return XCD ? (nb + kXcd - 1) / kXcd * kXcd : nb;
}
template <bool XCD>
__device__ __forceinline__ int64_t logical_block(int64_t nb) {
    const int64_t b = blockIdx.x;
    if (!XCD) return b;
    const int64_t per = (nb + kXcd - 1) / kXcd;
    return (b % kXcd) * per + b / kXcd;
}

// Row layout of everything held in position order (the D-ILU factors, the position-order
// operator and the Krylov vectors of the position-order solves).  Positions k are cut into
// tiles of kTile = 64 (the last tile holds F mod 64 positions when that is not zero) and
// a tile's rows are stored layer by layer: row (k, z) sits at pos_base(k) + z * pos_stride(k).
// A column kernel's wave (64 consecutive positions of one tile) then reads each array as
// one run of 64 L rows, and the layers of a face column lie 64 rows apart instead of F.
// With every column in its own layer plane 4 F bytes apart (q = z F + k), a colour launch
// touched K arrays x L planes per wave and streamed at 55-75 % of the rate it reaches with
// the tiles.  The map is a bijection onto [0, F L); for L = 1 it is the identity.
constexpr int kTileShift = 6;
constexpr int kTile = 1 << kTileShift;
__host__ __device__ __forceinline__ uint32_t pos_base(uint32_t k, uint32_t L) {
    return (((k >> kTileShift) * L) << kTileShift) + (k & (kTile - 1));
}
__host__ __device__ __forceinline__ uint32_t pos_stride(uint32_t k, uint32_t F) {
    return (k >> kTileShift) < (F >> kTileShift) ? (uint32_t)kTile : (F & (kTile - 1));
}
// the same in 64-bit arithmetic (kernels that do not assume 3 F L < 2^32)
__host__ __device__ __forceinline__ int64_t pos_base64(int64_t k, int64_t L) {
    return (((k >> kTileShift) * L) << kTileShift) + (k & (kTile - 1));
}
__host__ __device__ __forceinline__ int64_t pos_stride64(int64_t k, int64_t F) {
    return (k >> kTileShift) < (F >> kTileShift) ? (int64_t)kTile : (F & (kTile - 1));
}
__host__ __device__ __forceinline__ int64_t pos_row(int64_t k, int z, int64_t F, int64_t L) {
    return pos_base64(k, L) + (int64_t)z * pos_stride64(k, F);
}
// First position of the grid of a column launch over positions [k0, k1): k0 rounded down to
// its tile, so that every wave stays inside one tile (threads below k0 do no work).
__host__ __device__ __forceinline__ int64_t tile_floor(int64_t k0) { return k0 & ~(int64_t)(kTile - 1); }
// workgroups of B threads of a column launch over positions [k0, k1)
inline int64_t col_blocks(int64_t k0, int64_t k1, int B = kBlock) { return (k1 - tile_floor(k0) + B - 1) / B; }

struct AssembleArgs {
    int64_t F;
    int L;
    double dz, dt;
    // config
    int do_fixed_settling, do_sublimation, do_lateral_diff, enable_veg;
    int use_tanh_fetch, use_exp_fetch, use_R94_lambda;
    int use_PomLi_probability, z0_ustar_coupling, rouault_diffusion_coef;
    double settling_velocity, settling_scale, min_sd_trans, cutoff, snow_diffusion_const, z0_snow;
    // per-face values of z0_snow / settling_scale (a batch of parameter-ensemble members held
    // as one union), or null: the scalars above
    const double *z0_face, *settle_face;
    // geometry, edges j-major ([j*F + i])
    const double *nx, *ny, *elen, *area, *cz;
    const int32_t* neigh;
    const uint16_t* slots;  // 4-bit CSR positions of self / neighbour 0..2 in the lateral block
    const double *canopy, *lai, *stalk_n, *stalk_d;
    const uint8_t* is_water;
    // forcing
    const double *U_R, *U_2m, *vw_dir, *swe, *t, *rh, *snowdepthavg, *fetch, *p_snow_hours;
    // suspension system
    const int32_t* rp;
    double *val, *rhs, *x;
    // per-cell / per-face outputs
    double *u_z, *csubl, *Qsalt, *blowing_prob;
    uint8_t* salt_flag;
    double* fscal;  // [3F] c_salt, hs, u* (features of the learned warm start), may be null
    // fused preconditioner set-up (nullable)
    double *pc_lo, *pc_cp, *pc_im, *pc_invdiag;
    float *pc_lo32, *pc_cp32, *pc_im32;
};

struct PostArgs {
    int64_t F;
    int L;
    double dz, dt;
    const double *x, *u_z, *csubl;
    double *Qsusp, *Qsubl, *Qsubl_mass, *sum_subl;
};

struct DepositionArgs {
    int64_t F;
    const double *vw_dir, *nx, *ny, *elen, *area, *dep_w, *Qsusp, *Qsalt;
    const int32_t *neigh, *rp;
    const uint16_t* slots;
    double *val, *rhs, *x, *invdiag;
};

struct DriftArgs {
    int64_t F;
    double dt;
    const double *qdep, *swe;
    const uint8_t* salt_flag;
    double *drift_mass, *sum_drift, *more;
    // batch: face i belongs to member face_member[i], whose drift runs only when
    // member_on[member] != 0 (its deposition was present); null: every face
    const int32_t *face_member, *member_on;
};

__global__ void k_assemble(AssembleArgs a);
__global__ void k_post(PostArgs a);
__global__ void k_deposition(DepositionArgs a);
__global__ void k_drift(DriftArgs a);

// ---------------------------------------------------------------------------
// linear algebra (linalg.hip)
// ---------------------------------------------------------------------------
// y = A x for a CSR matrix, value type T (double / float), vector type T.
template <typename T>
void spmv(hipStream_t s, int64_t n, const int32_t* rp, const int32_t* col, const T* val, const T* x, T* y);
// r = b - A x  and per-block partial sums of r.r into partial[block]
template <typename T>
void residual(hipStream_t s, int64_t n, const int32_t* rp, const int32_t* col, const T* val, const T* x,
              const T* b, T* r, double* partial, int nblk);

// Structure-aware SpMV of the extruded prism-column operator in its CSR value
// layout: one thread per face walks the face's L rows; columns come from the
// neighbour list and the slot table instead of col[].  With b: r = b - A x and
// per-block partial sums of r.r (partial[block]).
template <typename T>
void spmv_col(hipStream_t s, int64_t F, int L, const int32_t* rp, const T* val, const int32_t* neigh,
              const uint16_t* slots, const T* x, const T* b, T* y, double* partial);
inline int64_t spmv_col_blocks(int64_t F) { return (F + kBlock - 1) / kBlock; }
// partial[block*(nv+1) + c] = sum over the block's rows of V_c . w (c < nv), w . w (c == nv)
template <typename T>
void mdot(hipStream_t s, int64_t n, const T* V, int nv, const T* w, double* partial, int nblk);
// out[c] = sum_b partial[b*stride + c] for c < ncols (fixed order, deterministic)
void reduce_cols(hipStream_t s, const double* partial, int nblk, int stride, int ncols, double* out);
// w -= sum_c h[c] V_c ; partial[block] = sum w.w (after the update)
template <typename T>
void gs_update(hipStream_t s, int64_t n, const T* V, int nv, const double* h, T* w, double* partial, int nblk);
// v = w * (1 / *nrm)   (nrm on device)
template <typename T>
void scale_by_inv(hipStream_t s, int64_t n, const T* w, const double* nrm, T* v);
// u = sum_c y[c] V_c  (y on device)
template <typename T>
void combine(hipStream_t s, int64_t n, const T* V, int nv, const double* y, T* u);
// max |b| per block into partial
void absmax(hipStream_t s, int64_t n, const double* b, double* partial, int nblk);
void reduce_max(hipStream_t s, const double* partial, int nblk, double* out);

// preconditioners (z = M^-1 r)
template <typename T>
void pc_column(hipStream_t s, int64_t F, int L, const T* lo, const T* cp, const T* im, const T* r, T* z);
template <typename T>
void pc_jacobi(hipStream_t s, int64_t n, const double* invdiag, const T* r, T* z);

// Multicolour D-ILU(0) (precond.hip).  Faces are greedily coloured (no two
// neighbours share a colour); rows are ordered (colour, layer, face).  In that
// order ILU(0) only modifies the diagonal, so the factors are the strict lower
// and upper parts of A plus d~_r = a_rr - sum_{k<r} a_rk a_kr / d~_k.
// Faces get a colour-major position k (colour by colour, ascending face id
// inside a colour); the factors live in position order (rows at pos_base(k) +
// z pos_stride(k), tiles of 64 positions) so that every colour phase reads and
// writes contiguous runs.
struct ColorOrder {
    int ncolor = 0;
    std::vector<int64_t> off;  // positions of colour c: [off[c], off[c+1])
    DevBuf<int32_t> faces;     // [F] position -> face id
    DevBuf<int32_t> nbp;       // [3F] j-major over positions: position of neighbour j, -1 if none
    DevBuf<uint8_t> rev;       // [3F] j-major over positions: edge of neighbour j that points back
    DevBuf<uint8_t> lowmask;   // [F] over positions: bit j = neighbour j has a lower colour
    DevBuf<uint8_t> edge;      // [F] over positions: slot j holds the neighbour of edge (edge >> 2j) & 3
    std::vector<int32_t> h_color;
    std::vector<int32_t> h_faces;  // host copy of faces: position -> face index
};
// Zeroes elements [from, n) of a buffer (the pad slots below); at set-up only.
template <typename T>
inline void zero_tail(DevBuf<T>& b, size_t from) {
    if (b.p && b.n > from) PB_HIP(hipMemset(b.p + from, 0, (b.n - from) * sizeof(T)));
}
// D-ILU factors in position order; row q = pos_base(k) + z pos_stride(k).  lat, y and t carry one zero pad slot
// past their last entry (lat[3N], y[N], t[N]): the split operator's sweeps read an absent
// neighbour there instead of branching round the load.
template <typename T>
struct DiluFactors {
    DevBuf<T> lat;   // [3N + 1] coupling to neighbour j (row's own lateral entries), zero pad
    DevBuf<T> vert;  // [2N] coupling below (0) / above (1)
    DevBuf<T> dinv;  // [N]  1 / d~
    DevBuf<T> diag;  // [N]  a_rr (with lat and vert: the operator itself in position order)
    DevBuf<T> E;     // [N]  a_rr - d~ (Eisenstat split operator)
    DevBuf<T> y;     // [N + 1]  intermediate of the apply, zero pad
    DevBuf<T> t;     // [N + 1]  intermediate of the split operator, zero pad
    void alloc(int64_t N) {
        lat.alloc(3 * N + 1);
        vert.alloc(2 * N);
        dinv.alloc(N);
        diag.alloc(N);
        E.alloc(N);
        y.alloc(N + 1);
        t.alloc(N + 1);
        zero_tail(lat, 3 * N);
        zero_tail(y, N);
        zero_tail(t, N);
    }
};
void build_color_order(ColorOrder& co, int64_t F, const int32_t* neigh_fm, hipStream_t s);
// the colouring alone (host): color[F], returns the number of colours
int face_colors(int64_t F, const int32_t* neigh_fm, int32_t* color);
template <typename T>
void dilu_factor(hipStream_t s, const ColorOrder& co, int64_t F, int L, const int32_t* rp, const T* val,
                 const uint16_t* slots, DiluFactors<T>& d);
// z = M^-1 r.  pos: r and z in position order (tiled rows), else in the reference order.
template <typename T>
void dilu_apply(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* r, T* z,
                bool pos = false, bool xcd = false);
// times the position-order apply with and without the XCD workgroup mapping (HIP events on the
// stream); true if the mapping is faster.  ms_out[2] receives both averages (ms), if given.
template <typename T>
bool tune_dilu_xcd(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* r, T* z,
                   float* ms_out);
// Split (Eisenstat) form of the D-ILU preconditioned operator, position order (nLayer <= 16):
// which = 0: out = Ahat in = D~ (D~ + L)^-1 A (D~ + U)^-1 in (one backward + one forward sweep);
// 1: out = D~ (D~ + L)^-1 in (left transform of a right-hand side);
// 2: out = (D~ + U)^-1 in (right transform back to the unknowns).  in and out must differ;
// for which = 2, out must hold N + 1 entries with out[N] = 0 (the sweeps gather from it).
// mode: bit 0 the XCD workgroup mapping, bit 1 the factor arrays loaded with the non-temporal hint,
// bit 2 (which = 0) the layer-parallel sweeps (a workgroup per 64 face columns, a wave per layer).
template <typename T>
void split_op(int which, hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* in, T* out,
              int mode);
// split_op(0) with per-block partial sums partial[2b] = dw.out, partial[2b+1] = out.out; returns
// the number of blocks b (at most F / kBlock + 9 per colour).  stop (may be NULL): the
// launches return at entry when *stop != 0 (a BiCGStab run that has finished).
template <typename T>
int64_t split_apply_dots(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* in, T* out,
                         const T* dw, double* partial, int mode, const double* stop = nullptr);
// Workgroup tables of a batch of independent systems held as one union (batch.hip):
// the union's faces are member-major, so each colour's positions are too, and the
// sweeps of colour c run as nblk[c] workgroups, tab[c][b] = {k0, k1, member, a} (faces
// [k0, k1) of one member, inside [a, a + kBlock) with a a multiple of kTile: thread t takes
// position a + t).  The forward launches of one application
// (colour 0 first, then 1 .. top) write their DOTS partial pairs at pair offset
// part_off[c] + b.
constexpr int kMaxColors = 8;
struct SweepTables {
    int ncolor = 0;
    DevBuf<int4> tab[kMaxColors];
    int64_t nblk[kMaxColors] = {};
    int64_t part_off[kMaxColors] = {};
    int64_t total = 0;  // partial pairs of one application
};
// split_op(0) of a batch: every workgroup returns at entry when stop[member * stop_stride]
// != 0 (that member's run has finished; stop may be NULL); with dw, the partial pairs
// (dw.out, out.out) of each workgroup go to partial[2 (part_off[c] + b)].  L <= 16.
template <typename T>
void split_apply_batch(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const SweepTables& st,
                       const T* in, T* out, const T* dw, double* partial, const double* stop, int stop_stride);
constexpr int kSplitModes = 8;
// times split_op(0) in each mode (HIP events on the stream) and returns the fastest;
// ms_out[kSplitModes] receives the averages (ms), if given.
template <typename T>
int tune_split_mode(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* w, T* o,
                    float* ms_out);
// FP32 copy of FP64 factors (the FP32 preconditioner and operator of the mixed solve)
void dilu_narrow(hipStream_t s, int64_t N, const DiluFactors<double>& a, DiluFactors<float>& b);
// y = b - A x with per-block partial sums of y.y, everything in position order, A taken
// from the factor arrays diag/lat/vert
template <typename T>
void pos_spmv(hipStream_t s, const ColorOrder& co, int64_t F, int L, const DiluFactors<T>& d, const T* x, const T* b,
              T* y, double* partial);
// y = A x in position order with launch shape v: v % 5 = 0 / 4 one thread per row (blocks of
// 256 / 128), 1 / 2 / 3 one thread per face column (blocks of 256 / 128 / 64; L <= 16); v >= 5
// with the XCD workgroup mapping
constexpr int kPosSpmvVariants = PBSM3D_SPMV_VARIANTS;
template <typename T>
void pos_spmv_variant(hipStream_t s, int v, const ColorOrder& co, int64_t F, int L, const DiluFactors<T>& d,
                      const T* x, T* y);
// the same SpMV with per-block partial sums partial[2b] = w.y, partial[2b+1] = y.y (FP64);
// returns the number of blocks b (at most n / 64 + 1024 for n = F * L, L >= 2)
template <typename T>
int64_t pos_spmv_dots(hipStream_t s, int v, const ColorOrder& co, int64_t F, int L, const DiluFactors<T>& d,
                      const T* x, T* y, const T* w, double* partial);
inline int64_t pos_spmv_dots_max_blocks(int64_t n) { return n / 64 + 1024; }
// times every launch shape on (x, y) with HIP events on the stream; returns the fastest
// (ms_out[v], if given, receives the average time of shape v, ms)
template <typename T>
int tune_pos_spmv(hipStream_t s, const ColorOrder& co, int64_t F, int L, const DiluFactors<T>& d, const T* x, T* y,
                  float* ms_out);
inline int64_t pos_spmv_blocks(int64_t F, int L) { return ((F + kBlock - 1) / kBlock) * L; }
// out[pos_row(k, z)] = in[z*F + face(k)]  /  out[z*F + face(k)] = in[pos_row(k, z)]
template <typename Ti, typename To>
void to_pos(hipStream_t s, const ColorOrder& co, int64_t F, int L, const Ti* in, To* out);
template <typename Ti, typename To>
void from_pos(hipStream_t s, const ColorOrder& co, int64_t F, int L, const Ti* in, To* out);

// BiCGStab vector updates (FP64 arithmetic on T storage)
// out = y - a x; partial[2*blk] = rh.out (if rh), partial[2*blk+1] = out.out
template <typename T>
void axpy_dots(hipStream_t s, int64_t n, double a, const T* x, const T* y, T* out, const T* rh, double* partial,
               int nblk);

// BiCGStab with its scalars kept on the device (slots of sc[]).  The convergence test is
// made on the device too: once sc[SC_STOP] is set every later BiCGStab kernel of the run
// returns at entry, so the host can queue iteration k+1 before it reads the result of k.
enum BicgSc {
    SC_RHO_PREV = 0, SC_RHO, SC_ALPHA, SC_OMEGA, SC_RHV, SC_SS, SC_ST, SC_TT, SC_RR, SC_FLAG, SC_STOP,
    SC_RHNN,      // rh.rh, the shadow residual's squared norm
    SC_RESTART,   // 1: the last stage restarted the shadow residual (the next p kernel sets rh = p = r)
    SC_NRESTART,  // restarts in this run
    SC_COUNT
};
// p = r + beta (p - omega v) with beta from the scalars of the previous step; after a
// restart of the shadow residual (sc[SC_RESTART]): p = r and rh = r
template <typename T>
void bicg_p_dev(hipStream_t s, int64_t n, const T* r, const T* v, const double* sc, T* p, T* rh);
// alpha = rho / (rh.v) and s = r - alpha v, partial[b] = s.s (nblk blocks).  dots: the
// operator's dots_nb per-block pairs (rh.v, v.v); every workgroup sums them in the same fixed
// order and workgroup 0 records rh.v and alpha in sc (no single-workgroup stage launch).
// partial must not alias dots.
template <typename T>
void bicg_s_dev(hipStream_t s, int64_t n, const T* v, const T* r, double* sc, T* s_out, double* partial, int nblk,
                const double* dots, int64_t dots_nb);
// omega from the operator's pairs (s.t, t.t) and k_bicg_s_dev's ss_nb partials of s.s (summed
// like alpha above), then x += alpha ph + omega sh; r = s - omega t; partial[2b] = rh.r,
// partial[2b+1] = r.r.  partial must not alias dots or ss_part.
template <typename T>
void bicg_xr_dev(hipStream_t s, int64_t n, const T* ph, const T* sh, const T* sv, const T* tv, const T* rh, double* sc,
                 T* x, T* r, double* partial, int nblk, const double* dots, int64_t dots_nb, const double* ss_part,
                 int64_t ss_nb);
// one block sums the pairs (rh.r, r.r) of nblk partials (stride doubles apart) in a fixed order:
// stage 0 starts a run (rho, rr; clears sc[SC_STOP]), stage 1 ends an iteration: new rho and
// rr, and sc[SC_STOP] set when a breakdown was flagged (sc[SC_FLAG] 1: rh.v = 0, 2: omega = 0,
// 3: non-finite residual), sqrt(rr) <= stop_abs or rr is not finite.  Stage 1 does nothing once
// sc[SC_STOP] is set.  host_slot (may be NULL): device-visible pinned host memory that
// receives sc[0 .. SC_COUNT) after the update.  restart_eps > 0: stage 1 restarts the shadow
// residual when |rh.r| < restart_eps ||rh|| ||r|| (a near breakdown of the Lanczos recurrence):
// rho = r.r, sc[SC_RESTART] = 1, and the next p kernel sets rh = p = r.
void bicg_stage(hipStream_t s, const double* partial, int64_t nblk, int stride, int stage, double* sc,
                double stop_abs = 0.0, double* host_slot = nullptr, double restart_eps = 0.0);

// conversions / vector updates
template <typename T>
void scale_host(hipStream_t s, int64_t n, const T* a, double alpha, T* b);  // b = alpha * a
void to_f32(hipStream_t s, int64_t n, const double* a, float* b);
void axpy_f32_into_f64(hipStream_t s, int64_t n, double alpha, const float* x, double* y);  // y += alpha*x
void scale_f64_to_f32_host(hipStream_t s, int64_t n, const double* r, double alpha, float* out);  // out = alpha*r
void axpy_dev_scaled_f32(hipStream_t s, int64_t n, const double* scale, const float* x, double* y);  // y += (*scale) x
void axpy_f64(hipStream_t s, int64_t n, double alpha, const double* x, double* y);
void copy_f64(hipStream_t s, int64_t n, const double* a, double* b);

}  // namespace pb
