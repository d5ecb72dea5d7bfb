// Per-face network of the learned warm start: 16 -> 128 -> 128 -> 128 -> 11,
// ReLU, BF16 operands with FP32 accumulation on the matrix cores
// (gfx950's v_mfma_f32_16x16x32_bf16), trained on the GPU by its own forward
// and backward kernels with Adam.
//
// Every layer is a GEMM with a fused epilogue.  Activations are stored
// [samples x features] row-major; for a layer with weights W [out x in]
//   forward   Z = A_in W^T + b,   A_out = relu(Z)        (M=samples, N=out, K=in)
//   backward  dW = dZ^T A_in,     db = colsum(dZ)        (M=out, N=in, K=samples)
//             dA_in = dZ W,       dZ_in = dA_in * [Z_in > 0]  (M=samples, N=in, K=out)
// The GEMM stages 64x32 A and 32x64 B tiles of BF16 in LDS, both k-contiguous
// (As[row][k], Bs[col][k]); each of the four wavefronts owns a 32x32 quarter of
// the 64x64 output tile, i.e. 2x2 MFMA 16x16 blocks, with the operand fragments
// read from LDS in the MFMA lane layout as one 16-byte load per lane
// (A: row = lane%16, k = 8*(lane/16)+j, j < 8; B: col = lane%16, same k;
// C: col = lane%16, row = 4*(lane/16)+j, j < 4).
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "mlp.hpp"

namespace pb {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32;

// 8 consecutive BF16 values (16 bytes, 16-byte aligned) as an MFMA operand fragment
__device__ __forceinline__ bf16x8 frag8(const uint16_t* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                    // round to nearest even
    return (uint16_t)(u >> 16);
}

struct GemmArgs {
    int M, N, K;
    const uint16_t* A;
    int lda, transA;
    const uint16_t* B;
    int ldb, transB;
    float* C;  // fp32 output (pre-activation), may be null
    int ldc;
    const float* bias;  // per output column, may be null
    const float* mask;  // relu' mask source (same shape as output, ld = ldc), may be null
    int relu;
    uint16_t* Cb;  // bf16 copy of the (activated) output, may be null
    int ldcb;
    float* slab;  // split-K: blockIdx.z writes its raw partial product to slab[z][M][N]
    int ksplit;   // K range per blockIdx.z
};

// One 16-byte chunk of a 64 x 32 operand tile into LDS T[64][TK + 8] (T[row][k]).
// kmajor = 0: element (row, k) at X[row * ld + k] (k contiguous; chunk = 8 k of one row);
// kmajor = 1: element (row, k) at X[k * ld + row] (row contiguous; chunk = 8 rows of one k).
// rows: the matrix's row count (M or N); r0 / k0: the tile origin; ke: end of the K range.
__device__ __forceinline__ void stage_chunk(const uint16_t* __restrict__ X, int ld, int kmajor, int rows, int r0,
                                            int k0, int ke, int tid, uint16_t (*T)[TK + 8]) {
    typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
    if (!kmajor) {
        const int r = tid >> 2, c = (tid & 3) * 8;  // 64 rows x 4 chunks
        const int gr = r0 + r, gk = k0 + c;
        u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
        const size_t off = (size_t)gr * ld + gk;
        if (gr < rows && gk + 8 <= ke && ((uintptr_t)(X + off) & 15) == 0) {
            v = *reinterpret_cast<const u16x8*>(X + off);
        } else if (gr < rows) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (gk + j < ke) v[j] = X[off + j];
        }
        *reinterpret_cast<u16x8*>(&T[r][c]) = v;
    } else {
        const int kk = tid >> 3, r = (tid & 7) * 8;  // 32 k x 8 chunks of 8 rows
        const int gk = k0 + kk, gr = r0 + r;
        u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
        const size_t off = (size_t)gk * ld + gr;
        if (gk < ke) {
            if (gr + 8 <= rows && ((uintptr_t)(X + off) & 15) == 0) {
                v = *reinterpret_cast<const u16x8*>(X + off);
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (gr + j < rows) v[j] = X[off + j];
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) T[r + j][kk] = v[j];
    }
}

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
    __shared__ __align__(16) uint16_t As[TM][TK + 8];
    __shared__ __align__(16) uint16_t Bs[TN][TK + 8];
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kb = blockIdx.z * g.ksplit;
    const int ke = min(g.K, kb + g.ksplit);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += TK) {
        // each thread stages one 8-element (16-byte) chunk of A and one of B: 64 x 32 = 256
        // chunks per operand.  A chunk runs along the operand's contiguous dimension (k for
        // A [M x K] / B [N x K], m or n for the transposed storage) and is loaded with one
        // vector load when it lies inside the matrix, element by element (zero-filled)
        // otherwise; the LDS layout (As[m][k], Bs[n][k]) is the same either way.
        stage_chunk(g.A, g.lda, g.transA, g.M, m0, k0, ke, tid, As);
        stage_chunk(g.B, g.ldb, g.transB ^ 1, g.N, n0, k0, ke, tid, Bs);
        __syncthreads();
#pragma unroll
        for (int tm = 0; tm < 2; ++tm) {
            const bf16x8 a = frag8(&As[wm + tm * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
            for (int tn = 0; tn < 2; ++tn) {
                const bf16x8 b = frag8(&Bs[wn + tn * 16 + (lane & 15)][8 * (lane >> 4)]);
                acc[tm][tn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[tm][tn], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int tm = 0; tm < 2; ++tm)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = m0 + wm + tm * 16 + 4 * (lane >> 4) + j;
                const int col = n0 + wn + tn * 16 + (lane & 15);
                if (row >= g.M || col >= g.N) continue;
                float v = acc[tm][tn][j];
                if (g.slab) {
                    g.slab[((size_t)blockIdx.z * g.M + row) * g.N + col] = v;
                    continue;
                }
                if (g.bias) v += g.bias[col];
                if (g.mask && !(g.mask[(size_t)row * g.ldc + col] > 0.f)) v = 0.f;
                if (g.C) g.C[(size_t)row * g.ldc + col] = v;
                if (g.relu) v = fmaxf(v, 0.f);
                if (g.Cb) g.Cb[(size_t)row * g.ldcb + col] = to_bf16(v);
            }
}

constexpr int kMaxSplits = 64;  // split-K factor of the weight-gradient GEMMs (K = batch rows)

// C[r][c] = sum over z = 0 .. splits-1, in that order, of slab[z][r][c] (deterministic split-K)
__global__ void k_sum_slabs(int M, int N, int splits, const float* __restrict__ slab, float* __restrict__ C,
                            int ldc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)M * N) return;
    const size_t mn = (size_t)M * N;
    float acc = 0.f;
    int z = 0;
    for (; z + 8 <= splits; z += 8) {  // eight loads in flight, added in slab order
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = slab[(size_t)(z + u) * mn + t];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; z < splits; ++z) acc += slab[(size_t)z * mn + t];
    C[(t / N) * ldc + t % N] = acc;
}

// Split-K GEMMs (splits > 1) need ws with room for splits * M * N floats; the
// partial products are summed in a fixed order, so results do not depend on
// scheduling (bias, masks and bf16 copies are not supported with splits).
void gemm(hipStream_t s, GemmArgs g, int splits = 1, float* ws = nullptr) {
    if (splits < 1) splits = 1;
    g.ksplit = ((g.K + splits - 1) / splits + TK - 1) / TK * TK;
    splits = (g.K + g.ksplit - 1) / g.ksplit;
    g.slab = splits > 1 ? ws : nullptr;
    if (splits > 1 && (!ws || g.bias || g.mask || g.relu || g.Cb || !g.C))
        throw Error(PBSM3D_EINVAL, "split-K GEMM needs a workspace and a plain FP32 output");
    dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM, splits);
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, g);
    if (splits > 1) {
        const int64_t mn = (int64_t)g.M * g.N;
        hipLaunchKernelGGL(k_sum_slabs, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, g.M, g.N, splits, ws,
                           g.C, g.ldc);
    }
}

// x (rows x 16, f32) -> normalised bf16 rows; optional row gather idx
__global__ void k_prep_inputs(int64_t n, const float* __restrict__ X, const int64_t* __restrict__ idx,
                              const float* __restrict__ mu, const float* __restrict__ sd, uint16_t* __restrict__ Xb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * MlpDims::IN) return;
    const int64_t r = t / MlpDims::IN;
    const int c = (int)(t % MlpDims::IN);
    const int64_t src = idx ? idx[r] : r;
    Xb[t] = to_bf16((X[src * MlpDims::IN + c] - mu[c]) / sd[c]);
}

// targets -> normalised f32 rows (OUTP wide, padding zero)
__global__ void k_prep_targets(int64_t n, const float* __restrict__ Y, const int64_t* __restrict__ idx,
                               const float* __restrict__ mu, const float* __restrict__ sd, float* __restrict__ T) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * MlpDims::OUTP) return;
    const int64_t r = t / MlpDims::OUTP;
    const int c = (int)(t % MlpDims::OUTP);
    const int64_t src = idx ? idx[r] : r;
    T[t] = c < MlpDims::OUT ? (Y[src * MlpDims::OUT + c] - mu[c]) / sd[c] : 0.f;
}

// dY = 2 (Yp - T) / (n * OUT) on the real outputs, and per-block sums of the squared error
__global__ void k_mse_grad(int64_t n, const float* __restrict__ Yp, const float* __restrict__ T,
                           uint16_t* __restrict__ dYb, float* __restrict__ dY, float* __restrict__ partial) {
    __shared__ float red[256];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float se = 0.f;
    if (t < n * MlpDims::OUTP) {
        const int c = (int)(t % MlpDims::OUTP);
        float g = 0.f;
        if (c < MlpDims::OUT) {
            const float d = Yp[t] - T[t];
            se = d * d;
            g = 2.f * d / (float)(n * MlpDims::OUT);
        }
        dYb[t] = to_bf16(g);
        if (dY) dY[t] = g;
    }
    red[threadIdx.x] = se;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Column sums of a bf16 matrix (rows x cols, cols dividing 256) into f32 (bias gradients),
// in two fixed-order passes: workgroup b sums rows [b R, (b + 1) R) with the 256 / cols
// threads of each column reading interleaved rows (a row's columns are read together, so
// the loads are coalesced; eight loads in flight per thread), then one workgroup per
// column adds the per-workgroup sums (each thread a strided subset, then a fixed-order
// block sum).
constexpr int kColsumBlocks = 512;
__global__ void __launch_bounds__(256) k_colsum_part(int64_t rows, int cols, int64_t R,
                                                     const uint16_t* __restrict__ D, float* __restrict__ part) {
    __shared__ float red[256];
    const int c = threadIdx.x % cols, lanes = 256 / cols, sub = threadIdx.x / cols;
    const int64_t r0 = (int64_t)blockIdx.x * R, r1 = min(rows, r0 + R);
    float s = 0.f;
    int64_t r = r0 + sub;
    for (; r + 7 * lanes < r1; r += 8 * lanes) {
        uint16_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = D[(r + u * lanes) * cols + c];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += __uint_as_float((uint32_t)v[u] << 16);
    }
    for (; r < r1; r += lanes) s += __uint_as_float((uint32_t)D[r * cols + c] << 16);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < cols) {
        float t = 0.f;
        for (int k = 0; k < lanes; ++k) t += red[k * cols + threadIdx.x];
        part[(int64_t)blockIdx.x * cols + threadIdx.x] = t;
    }
}
__global__ void __launch_bounds__(256) k_colsum_final(int nb, int cols, const float* __restrict__ part,
                                                      float* __restrict__ out) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float t = 0.f;
    for (int b = threadIdx.x; b < nb; b += 256) t += part[(int64_t)b * cols + c];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// *acc += scale * (sum of the partials, in a fixed order): one workgroup
__global__ void __launch_bounds__(256) k_loss_accum(int64_t nb, const float* __restrict__ partial, double scale,
                                                    double* __restrict__ acc) {
    __shared__ double red[256];
    double s = 0;
    for (int64_t k = threadIdx.x; k < nb; k += 256) s += partial[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *acc += red[0] * scale;
}

// Adam over all eight parameter tensors (four weight matrices, four bias vectors) in one
// launch: element t belongs to segment k with off[k] <= t < off[k + 1]
struct AdamSegs {
    float* p[8];
    float* m[8];
    float* v[8];
    const float* g[8];
    uint16_t* pb[8];
    const uint8_t* keep[8];
    int64_t off[9];
};
__global__ void k_adam_all(AdamSegs a, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.off[8]) return;
    int k = 0;
#pragma unroll
    for (int j = 1; j < 8; ++j) k += t >= a.off[j] ? 1 : 0;
    const int64_t i = t - a.off[k];
    if (a.keep[k] && !a.keep[k][i]) return;
    const float gt = a.g[k][i];
    const float mt = b1 * a.m[k][i] + (1.f - b1) * gt;
    const float vt = b2 * a.v[k][i] + (1.f - b2) * gt * gt;
    a.m[k][i] = mt;
    a.v[k][i] = vt;
    const float pt = a.p[k][i] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
    a.p[k][i] = pt;
    if (a.pb[k]) a.pb[k][i] = to_bf16(pt);
}

__global__ void k_to_bf16(int64_t n, const float* __restrict__ a, uint16_t* __restrict__ b) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = to_bf16(a[t]);
}

// de-normalise network outputs (OUTP wide) into OUT-wide rows
__global__ void k_denorm(int64_t n, const float* __restrict__ Yp, const float* __restrict__ mu,
                         const float* __restrict__ sd, float* __restrict__ Y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * MlpDims::OUT) return;
    const int64_t r = t / MlpDims::OUT;
    const int c = (int)(t % MlpDims::OUT);
    Y[t] = Yp[r * MlpDims::OUTP + c] * sd[c] + mu[c];
}

// Inference of the whole network in one launch: a workgroup walks 64-sample tiles
// (persistent grid) and a tile's activations ping-pong between two 64 x 128 BF16 LDS
// buffers, so nothing but the 16 input features and the 11 outputs of a sample touches
// HBM (the layer-by-layer path writes and re-reads 3 x 256 MB of activations at 1M
// faces).  The BF16 weights (73.7 KB) are read through the vector cache, where every
// workgroup of a CU shares them.  Eight waves; in a 128-wide layer wave w computes rows
// 16 (w % 4) .. +16 and columns 64 (w / 4) .. +64 as four 16 x 16 MFMA blocks, K in
// steps of 32 in increasing order -- the operand layout, accumulation order and epilogue
// (+ bias, ReLU, BF16 rounding) of k_gemm_bf16, so the output equals the layer-by-layer
// forward bit for bit.  LDS: 41.5 KB per workgroup.
constexpr int kInferRows = 64;
constexpr int kInferThreads = 512;
constexpr int kAs = MlpDims::H + 8;    // LDS row stride (BF16) of the activation tiles (16-byte rows)
constexpr int kAs0 = MlpDims::IN + 8;  // of the input tile

struct InferArgs {
    int64_t n;
    const float* X;        // [n x IN] raw features
    const uint16_t* W[4];  // BF16 [dout x din]
    const float* b[4];
    const float* norm;     // x_mean[IN], x_std[IN], y_mean[OUT], y_std[OUT]
    float* Y;              // [n x OUT] raw outputs
};

// one layer on a 64-row tile: in (LDS, row stride ins, K columns) x W^T (global BF16
// [dout x K]) + bias, dout = 128 (hidden) or 16 (output layer: waves 0-3 only); BF16
// (+ ReLU) into outb or FP32 into outf
__device__ __forceinline__ void infer_layer(const uint16_t* __restrict__ in, int ins, int K,
                                            const uint16_t* __restrict__ W, const float* __restrict__ bias, int dout,
                                            bool relu, uint16_t* __restrict__ outb, float* __restrict__ outf) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = 16 * (wave & 3);
    const int c0 = 64 * (wave >> 2);
    if (c0 >= dout) return;
    const int nblk = min(4, (dout - c0) / 16);  // 16-column MFMA blocks of this wave
    f32x4 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        // K = 16 (the input layer): the upper half of the 32-wide step is zero padding
        bf16x8 a = {};
        if (k0 + 8 * (lane >> 4) < K) a = frag8(&in[(r0 + (lane & 15)) * ins + k0 + 8 * (lane >> 4)]);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (b < nblk) {
                bf16x8 w = {};
                if (k0 + 8 * (lane >> 4) < K) w = frag8(&W[(c0 + 16 * b + (lane & 15)) * K + k0 + 8 * (lane >> 4)]);
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc[b], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        if (b < nblk) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = r0 + 4 * (lane >> 4) + j;
                const int col = c0 + 16 * b + (lane & 15);
                float v = acc[b][j] + bias[col];
                if (outf) {
                    outf[row * MlpDims::OUTP + col] = v;
                } else {
                    if (relu) v = fmaxf(v, 0.f);
                    outb[row * kAs + col] = to_bf16(v);
                }
            }
        }
    }
}

__global__ void __launch_bounds__(kInferThreads) k_mlp_infer(InferArgs p) {
    constexpr int H = MlpDims::H, IN = MlpDims::IN, OUT = MlpDims::OUT, OUTP = MlpDims::OUTP;
    __shared__ __align__(16) uint16_t xin[kInferRows * kAs0];
    __shared__ __align__(16) uint16_t act0[kInferRows * kAs];
    __shared__ __align__(16) uint16_t act1[kInferRows * kAs];
    __shared__ float yout[kInferRows * OUTP];
    const int tid = threadIdx.x;
    const float* xmu = p.norm;
    const float* xsd = p.norm + IN;
    const float* ymu = p.norm + 2 * IN;
    const float* ysd = p.norm + 2 * IN + OUT;
    const int64_t ntiles = (p.n + kInferRows - 1) / kInferRows;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t row0 = tile * kInferRows;
        __syncthreads();  // the previous tile's outputs are consumed
        for (int e = tid; e < kInferRows * IN; e += kInferThreads) {
            const int r = e / IN, c = e % IN;
            const int64_t g = row0 + r;
            xin[r * kAs0 + c] = g < p.n ? to_bf16((p.X[g * IN + c] - xmu[c]) / xsd[c]) : (uint16_t)0;
        }
        __syncthreads();
        infer_layer(xin, kAs0, IN, p.W[0], p.b[0], H, true, act0, nullptr);
        __syncthreads();
        infer_layer(act0, kAs, H, p.W[1], p.b[1], H, true, act1, nullptr);
        __syncthreads();
        infer_layer(act1, kAs, H, p.W[2], p.b[2], H, true, act0, nullptr);
        __syncthreads();
        infer_layer(act0, kAs, H, p.W[3], p.b[3], OUTP, false, nullptr, yout);
        __syncthreads();
        for (int e = tid; e < kInferRows * OUT; e += kInferThreads) {
            const int r = e / OUT, c = e % OUT;
            const int64_t g = row0 + r;
            if (g < p.n) p.Y[g * OUT + c] = yout[r * OUTP + c] * ysd[c] + ymu[c];
        }
    }
}

__global__ void k_sample_idx(int64_t nb, int64_t n, uint64_t key, int64_t* idx) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    uint64_t z = key + (uint64_t)t * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    idx[t] = (int64_t)(z % (uint64_t)n);
}

unsigned g1(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// layer l: W[l] is [dout(l) x din(l)] row-major
int MlpDims::din(int l) { return l == 0 ? IN : H; }
int MlpDims::dout(int l) { return l == 3 ? OUTP : H; }

Mlp::Mlp(int device_, uint64_t seed) : device(device_) {
    PB_HIP(hipSetDevice(device));
    PB_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int l = 0; l < 4; ++l) {
        const int di = MlpDims::din(l), dn = MlpDims::dout(l);
        std::vector<float> w((size_t)di * dn), b(dn, 0.f);
        const double sc = std::sqrt(2.0 / di);  // He initialisation
        for (int o = 0; o < dn; ++o)
            for (int i = 0; i < di; ++i) w[(size_t)o * di + i] = (l == 3 && o >= MlpDims::OUT) ? 0.f : (float)(nd(rng) * sc);
        W[l].alloc(w.size());
        Wb[l].alloc(w.size());
        bvec[l].alloc(dn);
        gW[l].alloc(w.size());
        gb[l].alloc(dn);
        mW[l].alloc(w.size());
        vW[l].alloc(w.size());
        mb[l].alloc(dn);
        vb[l].alloc(dn);
        PB_HIP(hipMemcpyAsync(W[l].p, w.data(), w.size() * 4, hipMemcpyHostToDevice, stream));
        PB_HIP(hipMemcpyAsync(bvec[l].p, b.data(), b.size() * 4, hipMemcpyHostToDevice, stream));
        for (auto* d : {&mW[l], &vW[l]}) PB_HIP(hipMemsetAsync(d->p, 0, d->bytes(), stream));
        for (auto* d : {&mb[l], &vb[l]}) PB_HIP(hipMemsetAsync(d->p, 0, d->bytes(), stream));
        hipLaunchKernelGGL(k_to_bf16, dim3(g1(w.size())), dim3(256), 0, stream, (int64_t)w.size(), W[l].p, Wb[l].p);
    }
    // output rows >= OUT stay zero: mask their updates
    std::vector<uint8_t> keep((size_t)MlpDims::OUTP * MlpDims::H, 1), keepb(MlpDims::OUTP, 1);
    for (int o = MlpDims::OUT; o < MlpDims::OUTP; ++o) {
        keepb[o] = 0;
        for (int i = 0; i < MlpDims::H; ++i) keep[(size_t)o * MlpDims::H + i] = 0;
    }
    keepW4.alloc(keep.size());
    keepb4.alloc(keepb.size());
    PB_HIP(hipMemcpyAsync(keepW4.p, keep.data(), keep.size(), hipMemcpyHostToDevice, stream));
    PB_HIP(hipMemcpyAsync(keepb4.p, keepb.data(), keepb.size(), hipMemcpyHostToDevice, stream));
    std::vector<float> one(32, 1.f), zero(32, 0.f);
    norm.alloc(2 * MlpDims::IN + 2 * MlpDims::OUT);
    std::vector<float> nrm(norm.n, 0.f);
    for (int c = 0; c < MlpDims::IN; ++c) nrm[MlpDims::IN + c] = 1.f;
    for (int c = 0; c < MlpDims::OUT; ++c) nrm[2 * MlpDims::IN + MlpDims::OUT + c] = 1.f;
    PB_HIP(hipMemcpyAsync(norm.p, nrm.data(), nrm.size() * 4, hipMemcpyHostToDevice, stream));
    PB_HIP(hipStreamSynchronize(stream));
}

Mlp::~Mlp() {
    if (stream) (void)hipStreamDestroy(stream);
}

void Mlp::ensure(int64_t n) {
    if (n <= cap) return;
    cap = n;
    Xb.alloc(n * MlpDims::IN);
    for (int l = 0; l < 3; ++l) {
        Z[l].alloc(n * MlpDims::H);
        Ab[l].alloc(n * MlpDims::H);
        dZb[l].alloc(n * MlpDims::H);
    }
    Yp.alloc(n * MlpDims::OUTP);
    Tn.alloc(n * MlpDims::OUTP);
    dYb.alloc(n * MlpDims::OUTP);
    idx.alloc(n);
    partial.alloc(g1(n * MlpDims::OUTP) + 1);
    if (!splitws.p) splitws.alloc((size_t)kMaxSplits * MlpDims::H * MlpDims::H);
    if (!colpart.p) colpart.alloc((size_t)kColsumBlocks * MlpDims::H);
    loss_dev.alloc(1);
}

const float* Mlp::xmu() const { return norm.p; }
const float* Mlp::xsd() const { return norm.p + MlpDims::IN; }
const float* Mlp::ymu() const { return norm.p + 2 * MlpDims::IN; }
const float* Mlp::ysd() const { return norm.p + 2 * MlpDims::IN + MlpDims::OUT; }

void Mlp::forward_prepared(int64_t n, bool keep_z) {
    hipStream_t s = stream;
    const uint16_t* in = Xb.p;
    for (int l = 0; l < 4; ++l) {
        GemmArgs g{};
        g.M = (int)n;
        g.N = MlpDims::dout(l);
        g.K = MlpDims::din(l);
        g.A = in;
        g.lda = MlpDims::din(l);
        g.B = Wb[l].p;
        g.ldb = MlpDims::din(l);
        g.transB = 1;
        g.bias = bvec[l].p;
        if (l < 3) {
            g.C = keep_z ? Z[l].p : nullptr;  // pre-activations: only the backward pass reads them
            g.ldc = MlpDims::H;
            g.relu = 1;
            g.Cb = Ab[l].p;
            g.ldcb = MlpDims::H;
            in = Ab[l].p;
        } else {
            g.C = Yp.p;
            g.ldc = MlpDims::OUTP;
        }
        gemm(s, g);
    }
}

double Mlp::backward_prepared(int64_t n, float* dY_f32, double* loss_acc) {
    hipStream_t s = stream;
    const unsigned nb = g1(n * MlpDims::OUTP);
    hipLaunchKernelGGL(k_mse_grad, dim3(nb), dim3(256), 0, s, n, Yp.p, Tn.p, dYb.p, dY_f32, partial.p);
    // weight gradients: K = n samples split into up to 64 slabs of >= 512 rows, so that the
    // 128 x 128 outputs (4 tiles) still give 256 workgroups at a batch of 32,768
    const int splits = (int)std::min<int64_t>(kMaxSplits, std::max<int64_t>(1, n / 512));
    const uint16_t* dcur = dYb.p;  // gradient w.r.t. this layer's pre-activation (bf16)
    for (int l = 3; l >= 0; --l) {
        const int di = MlpDims::din(l), dn = MlpDims::dout(l);
        const uint16_t* ain = l == 0 ? Xb.p : Ab[l - 1].p;
        GemmArgs gw{};  // gW = dZ^T A_in  (M=dout, N=din, K=n)
        gw.M = dn;
        gw.N = di;
        gw.K = (int)n;
        gw.A = dcur;
        gw.lda = dn;
        gw.transA = 1;
        gw.B = ain;
        gw.ldb = di;
        gw.C = gW[l].p;
        gw.ldc = di;
        gemm(s, gw, splits, splitws.p);
        {
            const int64_t R = std::max<int64_t>(64, (n + kColsumBlocks - 1) / kColsumBlocks);
            const int nbk = (int)((n + R - 1) / R);
            hipLaunchKernelGGL(k_colsum_part, dim3(nbk), dim3(256), 0, s, n, dn, R, dcur, colpart.p);
            hipLaunchKernelGGL(k_colsum_final, dim3(dn), dim3(256), 0, s, nbk, dn, colpart.p, gb[l].p);
        }
        if (l > 0) {  // dZ_{l-1} = (dZ W) * [Z_{l-1} > 0]   (M=n, N=din, K=dout)
            GemmArgs gd{};
            gd.M = (int)n;
            gd.N = di;
            gd.K = dn;
            gd.A = dcur;
            gd.lda = dn;
            gd.B = Wb[l].p;
            gd.ldb = di;
            gd.mask = Z[l - 1].p;
            gd.ldc = di;
            gd.Cb = dZb[l - 1].p;
            gd.ldcb = di;
            gemm(s, gd);
            dcur = dZb[l - 1].p;
        }
    }
    if (loss_acc) {  // training: the batch loss is added on the device, no host round trip per batch
        hipLaunchKernelGGL(k_loss_accum, dim3(1), dim3(256), 0, s, (int64_t)nb, partial.p, 1.0 / (double)(n * MlpDims::OUT),
                           loss_acc);
        return 0.0;
    }
    std::vector<float> hp(nb);
    PB_HIP(hipMemcpyAsync(hp.data(), partial.p, nb * sizeof(float), hipMemcpyDeviceToHost, s));
    PB_HIP(hipStreamSynchronize(s));
    double se = 0;
    for (float v : hp) se += v;
    return se / (double)(n * MlpDims::OUT);
}

void Mlp::adam_step(float lr) {
    ++t_adam;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)t_adam), c2 = 1.f - std::pow(b2, (float)t_adam);
    AdamSegs a{};
    a.off[0] = 0;
    for (int l = 0; l < 4; ++l) {
        const int64_t nw = (int64_t)MlpDims::din(l) * MlpDims::dout(l), nbias = MlpDims::dout(l);
        const int kw = 2 * l, kb = 2 * l + 1;
        a.p[kw] = W[l].p, a.m[kw] = mW[l].p, a.v[kw] = vW[l].p, a.g[kw] = gW[l].p, a.pb[kw] = Wb[l].p;
        a.keep[kw] = l == 3 ? keepW4.p : nullptr;
        a.off[kw + 1] = a.off[kw] + nw;
        a.p[kb] = bvec[l].p, a.m[kb] = mb[l].p, a.v[kb] = vb[l].p, a.g[kb] = gb[l].p, a.pb[kb] = nullptr;
        a.keep[kb] = l == 3 ? keepb4.p : nullptr;
        a.off[kb + 1] = a.off[kb] + nbias;
    }
    hipLaunchKernelGGL(k_adam_all, dim3(g1(a.off[8])), dim3(256), 0, stream, a, lr, b1, b2, eps, c1, c2);
}

void Mlp::prepare(const float* X, const float* Y, int64_t n, const int64_t* rows) {
    ensure(n);
    hipLaunchKernelGGL(k_prep_inputs, dim3(g1(n * MlpDims::IN)), dim3(256), 0, stream, n, X, rows, xmu(), xsd(), Xb.p);
    if (Y)
        hipLaunchKernelGGL(k_prep_targets, dim3(g1(n * MlpDims::OUTP)), dim3(256), 0, stream, n, Y, rows, ymu(), ysd(),
                           Tn.p);
}

double Mlp::train(const float* X, const float* Y, int64_t n, int epochs, int batch, float lr, uint64_t seed) {
    ensure(batch);
    const int64_t per_epoch = (n + batch - 1) / batch;
    double loss = 0;
    for (int e = 0; e < epochs; ++e) {
        PB_HIP(hipMemsetAsync(loss_dev.p, 0, sizeof(double), stream));
        for (int64_t b = 0; b < per_epoch; ++b) {
            const uint64_t key = seed * 0x9E3779B97F4A7C15ull + (uint64_t)(e * per_epoch + b) * 0xD1B54A32D192ED03ull;
            hipLaunchKernelGGL(k_sample_idx, dim3(g1(batch)), dim3(256), 0, stream, (int64_t)batch, n, key, idx.p);
            prepare(X, Y, batch, idx.p);
            forward_prepared(batch);
            backward_prepared(batch, nullptr, loss_dev.p);
            adam_step(lr);
        }
        double acc = 0;
        PB_HIP(hipMemcpyAsync(&acc, loss_dev.p, sizeof(double), hipMemcpyDeviceToHost, stream));
        PB_HIP(hipStreamSynchronize(stream));
        loss = acc / (double)per_epoch;
    }
    PB_HIP(hipStreamSynchronize(stream));
    return loss;
}

bool Mlp::fused_inference_ok() {
    if (fused_ok < 0) {
        int cus = 0;
        PB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        infer_blocks = 2 * (cus > 0 ? cus : 256);  // two workgroups per CU (104 VGPRs, 8 waves each)
        fused_ok = 1;
    }
    return fused_ok == 1;
}

void Mlp::predict_fused(const float* X, float* Y, int64_t n, hipStream_t s) {
    InferArgs a;
    a.n = n;
    a.X = X;
    for (int l = 0; l < 4; ++l) {
        a.W[l] = Wb[l].p;
        a.b[l] = bvec[l].p;
    }
    a.norm = norm.p;
    a.Y = Y;
    const int64_t tiles = (n + kInferRows - 1) / kInferRows;
    const unsigned grid = (unsigned)std::min<int64_t>(tiles, infer_blocks);
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(kInferThreads), 0, s, a);
    PB_HIP(hipGetLastError());
}

void Mlp::predict(const float* X, float* Y, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    if (fused_inference_ok()) {
        predict_fused(X, Y, n, s);
        return;
    }
    // layer by layer: runs on the caller's stream, chunked so that the workspace stays bounded
    predict_layers(X, Y, n, s);
}

void Mlp::predict_layers(const float* X, float* Y, int64_t n, hipStream_t s) {
    const hipStream_t own = stream;
    stream = s;
    const int64_t chunk = 1 << 20;
    for (int64_t o = 0; o < n; o += chunk) {
        const int64_t m = std::min(chunk, n - o);
        ensure(m);
        prepare(X + o * MlpDims::IN, nullptr, m, nullptr);
        forward_prepared(m, false);
        hipLaunchKernelGGL(k_denorm, dim3(g1(m * MlpDims::OUT)), dim3(256), 0, stream, m, Yp.p, ymu(), ysd(),
                           Y + o * MlpDims::OUT);
    }
    stream = own;
}

int64_t Mlp::param_count() {
    int64_t c = 0;
    for (int l = 0; l < 4; ++l) c += (int64_t)MlpDims::din(l) * MlpDims::dout(l) + MlpDims::dout(l);
    return c;
}

void Mlp::get_params(float* host, bool grads) {
    int64_t o = 0;
    for (int l = 0; l < 4; ++l) {
        const int64_t nw = (int64_t)MlpDims::din(l) * MlpDims::dout(l), nb = MlpDims::dout(l);
        PB_HIP(hipMemcpyAsync(host + o, grads ? gW[l].p : W[l].p, nw * 4, hipMemcpyDeviceToHost, stream));
        o += nw;
        PB_HIP(hipMemcpyAsync(host + o, grads ? gb[l].p : bvec[l].p, nb * 4, hipMemcpyDeviceToHost, stream));
        o += nb;
    }
    PB_HIP(hipStreamSynchronize(stream));
}

void Mlp::set_params(const float* host) {
    int64_t o = 0;
    for (int l = 0; l < 4; ++l) {
        const int64_t nw = (int64_t)MlpDims::din(l) * MlpDims::dout(l), nb = MlpDims::dout(l);
        PB_HIP(hipMemcpyAsync(W[l].p, host + o, nw * 4, hipMemcpyHostToDevice, stream));
        o += nw;
        PB_HIP(hipMemcpyAsync(bvec[l].p, host + o, nb * 4, hipMemcpyHostToDevice, stream));
        o += nb;
        hipLaunchKernelGGL(k_to_bf16, dim3(g1(nw)), dim3(256), 0, stream, nw, W[l].p, Wb[l].p);
    }
    PB_HIP(hipStreamSynchronize(stream));
}

void Mlp::set_normalization(const float* host) {
    PB_HIP(hipMemcpyAsync(norm.p, host, norm.bytes(), hipMemcpyHostToDevice, stream));
    PB_HIP(hipStreamSynchronize(stream));
}

double Mlp::gradients(const float* Xh, const float* Yh, int64_t n, float* grads_host) {
    DevBuf<float> X, Y;
    X.alloc(n * MlpDims::IN);
    Y.alloc(n * MlpDims::OUT);
    PB_HIP(hipMemcpyAsync(X.p, Xh, X.bytes(), hipMemcpyHostToDevice, stream));
    PB_HIP(hipMemcpyAsync(Y.p, Yh, Y.bytes(), hipMemcpyHostToDevice, stream));
    prepare(X.p, Y.p, n, nullptr);
    forward_prepared(n);
    const double loss = backward_prepared(n, nullptr);
    if (grads_host) get_params(grads_host, true);
    return loss;
}

}  // namespace pb
