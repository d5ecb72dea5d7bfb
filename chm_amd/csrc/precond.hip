// Multicolour D-ILU(0) for the extruded prism-column systems.
//
// Rows are taken in the order (face colour, layer, face).  A face's lateral
// neighbours all carry other colours, so within one colour the face columns
// are independent: a colour is one kernel launch in which a thread owns a
// face and walks its column -- bottom to top for the factorisation and the
// forward solve, top to bottom for the backward solve.  Two triangles sharing
// an edge have no third common neighbour except around a degree-3 vertex, so
// ILU(0) in this order changes only the diagonal (D-ILU):
//   d~_r = a_rr - sum_{k lower} a_rk a_kr / d~_k
//   M = (D~ + L) D~^-1 (D~ + U),  L / U the strict lower / upper parts of A.
//
// Layout: faces are given colour-major positions k, and the factor data
// (lateral and vertical couplings, 1/d~, the intermediate y) are stored at
// row pos_base(k) + z pos_stride(k): tiles of 64 positions, each tile's rows
// layer by layer, so a wave of a colour phase touches one run of 64 L rows
// per array.
// The factorisation gathers each row from the CSR once per step; the apply
// gathers r and scatters z once each.  The same kernels serve the 2-D
// deposition system (one layer).
#include <algorithm>
#include <vector>

#include "kernels.hpp"

namespace pb {
namespace {

__device__ __forceinline__ int nib(uint32_t slot, int k) { return (slot >> (4 * k)) & 0xF; }

// coefficient load; NT: with the non-temporal hint (read once per launch, so they need
// not displace the neighbour values the gathers reuse from L2)
// Operand of the gathers for an absent or masked neighbour: 0 * 0 is added instead of
// branching round the load, so every load of a column is unconditional and the compiler
// issues them together (a branch per neighbour put a full memory round trip per gather).
__device__ const double kZeroOperand[2] = {0.0, 0.0};
template <typename T>
__device__ __forceinline__ const T* zero_ptr() {
    return reinterpret_cast<const T*>(kZeroOperand);
}

template <bool NT, typename T>
__device__ __forceinline__ T ldc(const T* p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_dilu_factor(int64_t F, int L, int64_t k0, int64_t k1,
                                                        const int32_t* __restrict__ faces,
                                                        const int32_t* __restrict__ nbp,
                                                        const uint8_t* __restrict__ rev,
                                                        const uint8_t* __restrict__ lowmask,
                                                        const uint8_t* __restrict__ edge,
                                                        const int32_t* __restrict__ rp, const T* __restrict__ val,
                                                        const uint16_t* __restrict__ slots, T* __restrict__ lat,
                                                        T* __restrict__ vert, T* __restrict__ dinv,
                                                        T* __restrict__ diag, T* __restrict__ E) {
    const int64_t k = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= k1) return;
    const int64_t N = F * L;
    const int64_t i = faces[k];
    const uint32_t si = slots[i];
    const uint32_t lm = lowmask[k];
    const uint32_t ek = edge[k];  // slot f holds the neighbour of edge (ek >> 2f) & 3
    int32_t np_[3], nf[3];
    uint32_t sn[3];
    int rv[3];
    int64_t nq[3], ns[3];  // rows of the lower neighbours' columns: nq + z ns
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        np_[f] = nbp[f * F + k];
        const bool low = (lm >> f) & 1u;
        nf[f] = low ? faces[np_[f]] : 0;
        sn[f] = low ? slots[nf[f]] : 0u;
        rv[f] = low ? rev[f * F + k] : 0;
        nq[f] = pos_base64(max(np_[f], 0), L);  // rows of an absent neighbour: never read
        ns[f] = pos_stride64(max(np_[f], 0), F);
    }
    const int64_t qb = pos_base64(k, L), qs = pos_stride64(k, F);
    T dprev = 0;
    for (int z = 0; z < L; ++z) {
        const int64_t r = (int64_t)z * F + i;
        const int64_t q = qb + (int64_t)z * qs;
        const int64_t base = (int64_t)rp[r] + (z > 0 ? 1 : 0);
        T d = val[base + nib(si, 0)];
        diag[q] = d;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const T a_in = np_[f] >= 0 ? val[base + nib(si, (int)((ek >> (2 * f)) & 3u) + 1)] : T(0);
            lat[f * N + q] = a_in;
            if ((lm >> f) & 1u) {
                const int64_t rn = (int64_t)z * F + nf[f];
                const int64_t bn = (int64_t)rp[rn] + (z > 0 ? 1 : 0);
                const T a_ni = val[bn + nib(sn[f], rv[f] + 1)];
                d -= a_in * a_ni * dinv[nq[f] + (int64_t)z * ns[f]];
            }
        }
        const T a_ib = z > 0 ? val[rp[r]] : T(0);            // (r, r-F): first entry of row r
        const T a_ia = z < L - 1 ? val[rp[r + 1] - 1] : T(0);  // (r, r+F): last entry of row r
        vert[q] = a_ib;
        vert[N + q] = a_ia;
        if (z > 0) {
            const T a_bi = val[rp[r - F + 1] - 1];  // (r-F, r): last entry of row r-F
            d -= a_ib * a_bi * dprev;
        }
        const T di = d != T(0) ? T(1) / d : T(1);
        dinv[q] = di;
        E[q] = diag[q] - (d != T(0) ? d : T(1));  // D - D~ (Eisenstat split operator)
        dprev = di;
    }
}

// Forward sweep of one colour: y = D~^-1 (r - L y).  POS: r is in position
// order (tiled rows) rather than in the reference order (z*F + face).
// Neighbours of a face carry other colours, so the y values this launch reads
// (ynb, at neighbour positions) and writes (y, own positions) are disjoint
// elements of one array; the two restrict pointers let the loads run ahead
// of the stores.  LM > 0: the column (L <= LM) is fully unrolled -- first all
// terms that do not depend on the vertical recurrence, then the recurrence.
// TOP: this is the highest colour, whose faces have no higher-colour
// neighbours, so its backward sweep (z = y - D~^-1 U z, U here only the
// coupling above) runs in the same thread right after: z goes to zout (and, in
// the reference order, also back into y, where lower colours read it).
// XCD: workgroups are remapped so each XCD works on one contiguous range.
template <typename T, bool POS, int LM, bool TOP, bool XCD>
__global__ void __launch_bounds__(kBlock) k_dilu_fwd(int64_t F, int L, int64_t k0, int64_t k1,
                                                     const int32_t* __restrict__ faces, const int32_t* __restrict__ nbp,
                                                     const uint8_t* __restrict__ lowmask, const T* __restrict__ lat,
                                                     const T* __restrict__ vert, const T* __restrict__ dinv,
                                                     const T* __restrict__ rin, const T* __restrict__ ynb,
                                                     T* __restrict__ y, T* __restrict__ zout) {
    const int64_t kb = tile_floor(k0);
    const int64_t k = kb + logical_block<XCD>((k1 - kb + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
    if (k < k0 || k >= k1) return;
    const int64_t N = F * L;
    const int64_t i = POS ? k : (int64_t)faces[k];
    const uint32_t lm = lowmask[k];
    const int64_t qb = pos_base64(k, L), qs = pos_stride64(k, F);
    int32_t np_[3];
    int64_t nq[3], ns[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        np_[f] = ((lm >> f) & 1u) ? nbp[f * F + k] : -1;
        nq[f] = pos_base64(max(np_[f], 0), L);  // branch-free: the loads stay together
        ns[f] = pos_stride64(max(np_[f], 0), F);
    }
    // row z of r / z: position order (POS) or the reference order
    auto io = [&](int z) { return POS ? qb + (int64_t)z * qs : (int64_t)z * F + i; };
    auto independent = [&](int z) {
        const int64_t q = qb + (int64_t)z * qs;
        T acc = rin[io(z)];
#pragma unroll
        for (int f = 0; f < 3; ++f)
        {
            const bool on = np_[f] >= 0;
            acc -= *(on ? &lat[f * N + q] : zero_ptr<T>()) * *(on ? &ynb[nq[f] + (int64_t)z * ns[f]] : zero_ptr<T>());
        }
        return acc;
    };
    if constexpr (LM > 0) {
        T sv[LM], vv[LM], dv[LM], va[LM];
#pragma unroll
        for (int z = 0; z < LM; ++z)
            if (z < L) {
                const int64_t q = qb + (int64_t)z * qs;
                sv[z] = independent(z);
                vv[z] = z > 0 ? vert[q] : T(0);
                dv[z] = dinv[q];
                if (TOP) va[z] = z < L - 1 ? vert[N + q] : T(0);
            }
        T prev = 0;
#pragma unroll
        for (int z = 0; z < LM; ++z)
            if (z < L) {
                prev = (sv[z] - vv[z] * prev) * dv[z];
                if (TOP)
                    sv[z] = prev;
                else
                    y[qb + (int64_t)z * qs] = prev;
            }
        if (TOP) {
            T next = 0;
#pragma unroll
            for (int z = LM - 1; z >= 0; --z)
                if (z < L) {
                    next = sv[z] - dv[z] * (T(0) + va[z] * next);
                    if (!POS) y[qb + (int64_t)z * qs] = next;
                    zout[io(z)] = next;
                }
        }
    } else {
        T prev = 0;
        for (int z = 0; z < L; ++z) {
            const int64_t q = qb + (int64_t)z * qs;
            T acc = independent(z);
            if (z > 0) acc -= vert[q] * prev;
            prev = acc * dinv[q];
            y[q] = prev;
        }
        if (TOP) {
            T next = 0;
            for (int z = L - 1; z >= 0; --z) {
                const int64_t q = qb + (int64_t)z * qs;
                T acc = 0;
                if (z < L - 1) acc += vert[N + q] * next;
                next = y[q] - dinv[q] * acc;
                if (!POS) y[q] = next;
                zout[io(z)] = next;
            }
        }
    }
}

// Backward sweep of one colour: z = y - D~^-1 U z.  Reference order: y is
// updated in place (higher-colour neighbours are read from it through done)
// and z is scattered to z*F + face.  POS: z is in position order, the
// neighbours are read from z itself (done) and y is only read.  LM and XCD as
// in the forward sweep.
template <typename T, bool POS, int LM, bool XCD>
__global__ void __launch_bounds__(kBlock) k_dilu_bwd(int64_t F, int L, int64_t k0, int64_t k1,
                                                     const int32_t* __restrict__ faces, const int32_t* __restrict__ nbp,
                                                     const uint8_t* __restrict__ lowmask, const T* __restrict__ lat,
                                                     const T* __restrict__ vert, const T* __restrict__ dinv,
                                                     T* __restrict__ y, const T* __restrict__ done,
                                                     T* __restrict__ zout) {
    const int64_t kb = tile_floor(k0);
    const int64_t k = kb + logical_block<XCD>((k1 - kb + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
    if (k < k0 || k >= k1) return;
    const int64_t N = F * L;
    const int64_t i = POS ? k : (int64_t)faces[k];
    const uint32_t lm = lowmask[k];
    const int64_t qb = pos_base64(k, L), qs = pos_stride64(k, F);
    int32_t np_[3];
    int64_t nq[3], ns[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const int32_t n = nbp[f * F + k];
        np_[f] = (n >= 0 && !((lm >> f) & 1u)) ? n : -1;
        nq[f] = pos_base64(max(np_[f], 0), L);  // branch-free: the loads stay together
        ns[f] = pos_stride64(max(np_[f], 0), F);
    }
    auto io = [&](int z) { return POS ? qb + (int64_t)z * qs : (int64_t)z * F + i; };
    auto independent = [&](int z) {
        const int64_t q = qb + (int64_t)z * qs;
        T acc = 0;
#pragma unroll
        for (int f = 0; f < 3; ++f)
        {
            const bool on = np_[f] >= 0;
            acc += *(on ? &lat[f * N + q] : zero_ptr<T>()) * *(on ? &done[nq[f] + (int64_t)z * ns[f]] : zero_ptr<T>());
        }
        return acc;
    };
    if constexpr (LM > 0) {
        T av[LM], vv[LM], dv[LM], yv[LM];
#pragma unroll
        for (int z = 0; z < LM; ++z)
            if (z < L) {
                const int64_t q = qb + (int64_t)z * qs;
                av[z] = independent(z);
                vv[z] = z < L - 1 ? vert[N + q] : T(0);
                dv[z] = dinv[q];
                yv[z] = y[q];
            }
        T next = 0;
#pragma unroll
        for (int z = LM - 1; z >= 0; --z)
            if (z < L) {
                next = yv[z] - dv[z] * (av[z] + vv[z] * next);
                if (!POS) y[qb + (int64_t)z * qs] = next;
                zout[io(z)] = next;
            }
    } else {
        T next = 0;
        for (int z = L - 1; z >= 0; --z) {
            const int64_t q = qb + (int64_t)z * qs;
            T acc = independent(z);
            if (z < L - 1) acc += vert[N + q] * next;
            next = y[q] - dinv[q] * acc;
            if (!POS) y[q] = next;
            zout[io(z)] = next;
        }
    }
}

// ---------------------------------------------------------------------------
// Split (Eisenstat) form of the D-ILU preconditioned operator, position order.
// With A = (D~ + L) + (D~ + U) - (2 D~ - D) and M = (D~ + L) D~^-1 (D~ + U):
//   Ahat = D~ (D~ + L)^-1 A (D~ + U)^-1,   Ahat w = w + E t - L v,
// where E = D - D~, t = (D~ + U)^-1 w (backward sweep) and v solves
// (D~ + L) v = w - (2 D~ - D) t, i.e. v = D~^-1 (w + E t - L v) - t (forward
// sweep).  One backward and one forward sweep take the place of
// "precondition, then multiply by A".  Ahat x^ = b^, b^ = D~ (D~ + L)^-1 b,
// x = (D~ + U)^-1 x^ is the same system (Eisenstat 1981).
// ---------------------------------------------------------------------------

// t = D~^-1 (w - U t) for one colour; tnb reads t at the (higher-colour)
// neighbour positions, t is written at the colour's own positions.  LAT = false: the
// highest colour, whose neighbours all carry lower colours, so U has only the coupling
// above and the column reads no neighbour data at all.
template <typename T, int LM, bool EX, bool NT, bool LAT = true>
__device__ __forceinline__ void eis_bwd_column(int64_t F, int Lrt, int64_t k, const int32_t* __restrict__ nbp,
                                               const uint8_t* __restrict__ lowmask, const T* __restrict__ lat,
                                               const T* __restrict__ vert, const T* __restrict__ dinv,
                                               const T* __restrict__ w, const T* __restrict__ tnb,
                                               T* __restrict__ t) {
    // 32-bit positions (the host checks 3 N + 1 < 2^32): every load is base + 32-bit offset
    // from one uniform base.  An absent or masked neighbour reads the zero pad slots lat[3N]
    // and tnb[N] instead of branching round the load.
    const int L = EX ? LM : Lrt;  // EX: the layer count is the unroll bound, no guards
    const uint32_t Fu = (uint32_t)F, ku = (uint32_t)k;
    const uint32_t N = Fu * (uint32_t)L;
    const uint32_t qb = pos_base(ku, (uint32_t)L), qs = pos_stride(ku, Fu);
    const uint32_t lm = LAT ? lowmask[k] : 7u;
    int32_t np_[3];
    uint32_t nq[3], ns[3];  // rows of the neighbour columns: nq + z ns
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const int32_t n = LAT ? nbp[f * F + k] : -1;
        np_[f] = (n >= 0 && !((lm >> f) & 1u)) ? n : -1;
        // branch-free (an absent neighbour's rows are never read): a branch here would
        // hold each index load behind the previous one
        nq[f] = pos_base((uint32_t)max(n, 0), (uint32_t)L);
        ns[f] = pos_stride((uint32_t)max(n, 0), Fu);
    }
    T av[LM], va[LM], dv[LM], wv[LM];
#pragma unroll
    for (int z = 0; z < LM; ++z)
        if (z < L) {
            const uint32_t q = qb + (uint32_t)z * qs;
            T acc = 0;
            if (LAT) {
#pragma unroll
                for (int f = 0; f < 3; ++f)
                {
                    const bool on = np_[f] >= 0;
                    acc += ldc<NT>(&lat[on ? (uint32_t)f * N + q : 3 * N]) *
                           tnb[on ? nq[f] + (uint32_t)z * ns[f] : N];
                }
            }
            av[z] = acc;
            va[z] = z < L - 1 ? ldc<NT>(&vert[N + q]) : T(0);
            dv[z] = ldc<NT>(&dinv[q]);
            wv[z] = w[q];
        }
    T next = 0;
#pragma unroll
    for (int z = LM - 1; z >= 0; --z)
        if (z < L) {
            next = dv[z] * (wv[z] - (av[z] + va[z] * next));
            t[qb + (uint32_t)z * qs] = next;
        }
}

template <typename T, int LM, bool EX, bool XCD, bool NT, bool LAT = true>
__global__ void __launch_bounds__(kBlock) k_eis_bwd(int64_t F, int Lrt, int64_t k0, int64_t k1,
                                                    const int32_t* __restrict__ nbp,
                                                    const uint8_t* __restrict__ lowmask, const T* __restrict__ lat,
                                                    const T* __restrict__ vert, const T* __restrict__ dinv,
                                                    const T* __restrict__ w, const T* __restrict__ tnb,
                                                    T* __restrict__ t, const double* __restrict__ stop) {
    if (stop && *stop != 0.0) return;  // a finished BiCGStab run (uniform over the grid)
    const int64_t kb = tile_floor(k0);
    const int64_t k = kb + logical_block<XCD>((k1 - kb + kBlock - 1) / kBlock) * kBlock + threadIdx.x;
    if (k < k0 || k >= k1) return;
    eis_bwd_column<T, LM, EX, NT, LAT>(F, Lrt, k, nbp, lowmask, lat, vert, dinv, w, tnb, t);
}

// Forward sweep of one colour: acc = L v (lower neighbours from vnb, the cell
// below from the column), o = w + E t - acc, v = D~^-1 o - t (WITH_T), or
// o = w - acc, v = D~^-1 o (without t: the left transform b^ = D~ (D~ + L)^-1 b).
// FUSE: colour 0 (no lower neighbours), whose backward sweep runs first in the
// same thread and whose t stays in registers.
// PRE_DW: the dot vector is loaded with the rest of the column (more loads in flight, ten
// more VGPRs) rather than in the final loop.  SV = false: v is not stored (the top colour,
// whose v no higher colour gathers)
template <typename T, int LM, bool EX, bool WITH_T, bool FUSE, bool DOTS, bool NT, bool PRE_DW = true, bool SV = true>
__device__ __forceinline__ void eis_fwd_column(int64_t F, int Lrt, int64_t k, const int32_t* __restrict__ nbp,
                                               const uint8_t* __restrict__ lowmask, const T* __restrict__ lat,
                                               const T* __restrict__ vert, const T* __restrict__ dinv,
                                               const T* __restrict__ E, const T* __restrict__ w,
                                               const T* __restrict__ t, const T* __restrict__ tnb,
                                               const T* __restrict__ vnb, T* __restrict__ v, T* __restrict__ o,
                                               const T* __restrict__ dw, double& d1, double& d2) {
    // 32-bit positions and the zero pad slots lat[3N], vnb[N], tnb[N] (see eis_bwd_column)
    const int L = EX ? LM : Lrt;  // EX: the layer count is the unroll bound, no guards
    const uint32_t Fu = (uint32_t)F, ku = (uint32_t)k;
    const uint32_t N = Fu * (uint32_t)L;
    const uint32_t qb = pos_base(ku, (uint32_t)L), qs = pos_stride(ku, Fu);
    const uint32_t lm = lowmask[k];
    int32_t lo[3], up[3];
    uint32_t nq[3], ns[3];  // rows of the neighbour columns (lower or upper): nq + z ns
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const int32_t n = nbp[f * F + k];
        const bool low = (lm >> f) & 1u;
        lo[f] = (n >= 0 && low) ? n : -1;
        up[f] = (n >= 0 && !low) ? n : -1;
        nq[f] = pos_base((uint32_t)max(n, 0), (uint32_t)L);  // branch-free, as in eis_bwd_column
        ns[f] = pos_stride((uint32_t)max(n, 0), Fu);
    }
    T wv[LM], dv[LM], vb[LM], lv[LM], tv[LM], ev[LM], dwv[LM];
#pragma unroll
    for (int z = 0; z < LM; ++z)
        if (z < L) {
            const uint32_t q = qb + (uint32_t)z * qs;
            wv[z] = w[q];
            if (DOTS && PRE_DW) dwv[z] = dw[q];
            dv[z] = ldc<NT>(&dinv[q]);
            vb[z] = z > 0 ? ldc<NT>(&vert[q]) : T(0);
            T acc = 0;
            if (!FUSE) {
#pragma unroll
                for (int f = 0; f < 3; ++f)
                {
                    const bool on = lo[f] >= 0;
                    acc += ldc<NT>(&lat[on ? (uint32_t)f * N + q : 3 * N]) *
                           vnb[on ? nq[f] + (uint32_t)z * ns[f] : N];
                }
            }
            lv[z] = acc;
            if (WITH_T) {
                ev[z] = ldc<NT>(&E[q]);
                if (!FUSE) tv[z] = t[q];
            }
        }
    if constexpr (FUSE) {
        // backward sweep of this column first: t = D~^-1 (w - U t), all neighbours upper
        T av[LM], va[LM];
#pragma unroll
        for (int z = 0; z < LM; ++z)
            if (z < L) {
                const uint32_t q = qb + (uint32_t)z * qs;
                T acc = 0;
#pragma unroll
                for (int f = 0; f < 3; ++f)
                {
                    const bool on = up[f] >= 0;
                    acc += ldc<NT>(&lat[on ? (uint32_t)f * N + q : 3 * N]) *
                           tnb[on ? nq[f] + (uint32_t)z * ns[f] : N];
                }
                av[z] = acc;
                va[z] = z < L - 1 ? ldc<NT>(&vert[N + q]) : T(0);
            }
        T next = 0;
#pragma unroll
        for (int z = LM - 1; z >= 0; --z)
            if (z < L) {
                next = dv[z] * (wv[z] - (av[z] + va[z] * next));
                tv[z] = next;
            }
    }
    T prev = 0;
#pragma unroll
    for (int z = 0; z < LM; ++z)
        if (z < L) {
            const uint32_t q = qb + (uint32_t)z * qs;
            const T acc = lv[z] + vb[z] * prev;
            T oz, vz;
            if (WITH_T) {
                oz = wv[z] + ev[z] * tv[z] - acc;
                vz = dv[z] * oz - tv[z];
            } else {
                oz = wv[z] - acc;
                vz = dv[z] * oz;
            }
            if (SV) v[q] = vz;
            o[q] = oz;
            prev = vz;
            if (DOTS) {
                d1 = fma((double)(PRE_DW ? dwv[z] : dw[q]), (double)oz, d1);
                d2 = fma((double)oz, (double)oz, d2);
            }
        }
}

// DOTS: per-block partial sums (dw . o, o . o) into partial[2 * logical block] (FP64).
template <typename T, int LM, bool EX, bool XCD, bool WITH_T, bool FUSE, bool DOTS, bool NT, bool SV = true>
__global__ void __launch_bounds__(kBlock) k_eis_fwd(int64_t F, int L, int64_t k0, int64_t k1,
                                                    const int32_t* __restrict__ nbp,
                                                    const uint8_t* __restrict__ lowmask, const T* __restrict__ lat,
                                                    const T* __restrict__ vert, const T* __restrict__ dinv,
                                                    const T* __restrict__ E, const T* __restrict__ w,
                                                    const T* __restrict__ t, const T* __restrict__ tnb,
                                                    const T* __restrict__ vnb, T* __restrict__ v,
                                                    T* __restrict__ o, const T* __restrict__ dw,
                                                    double* __restrict__ partial, const double* __restrict__ stop) {
    if (stop && *stop != 0.0) return;  // a finished BiCGStab run (uniform over the grid)
    const int64_t kb = tile_floor(k0);
    const int64_t lb = logical_block<XCD>((k1 - kb + kBlock - 1) / kBlock);
    const int64_t k = kb + lb * kBlock + threadIdx.x;
    double d1 = 0, d2 = 0;
    if (k >= k0 && k < k1)
        eis_fwd_column<T, LM, EX, WITH_T, FUSE, DOTS, NT, true, SV>(F, L, k, nbp, lowmask, lat, vert, dinv, E, w, t,
                                                                    tnb, vnb, v, o, dw, d1, d2);
    if constexpr (DOTS) {
        // partials indexed by the logical block: the fixed-order sum sees the same faces in
        // the same order with and without the XCD workgroup mapping (blocks past the colour
        // add zeros), so the tuner's choice of mapping does not change the iterates
        __shared__ double s_red[kBlock / 64];
        const double a = block_sum(d1, s_red);
        const double b = block_sum(d2, s_red);
        if (threadIdx.x == 0) {
            partial[2 * lb] = a;
            partial[2 * lb + 1] = b;
        }
    }
}


// Table-driven forms for a batch of independent systems held as one union (members'
// faces consecutive, so each colour's positions are member-major): workgroup b takes
// faces [tab[b].x, tab[b].y) of one colour of member tab[b].z and returns at once when
// that member's run has stopped (stop[member * stop_stride] != 0).  No workgroup spans
// two members, so the DOTS partials of a workgroup belong to one member.  Thread t of the
// workgroup takes position tab[b].w + t (tab[b].w a tile boundary, so each wave stays in
// one tile) when it lies in [tab[b].x, tab[b].y).
template <typename T, int LM, bool EX, bool NT, bool LAT = true>
__global__ void __launch_bounds__(kBlock) k_eis_bwd_tab(int64_t F, int Lrt, const int4* __restrict__ tab,
                                                        const int32_t* __restrict__ nbp,
                                                        const uint8_t* __restrict__ lowmask, const T* __restrict__ lat,
                                                        const T* __restrict__ vert, const T* __restrict__ dinv,
                                                        const T* __restrict__ w, const T* __restrict__ tnb,
                                                        T* __restrict__ t, const double* __restrict__ stop,
                                                        int stop_stride) {
    const int4 e = tab[blockIdx.x];
    if (stop && stop[(int64_t)e.z * stop_stride] != 0.0) return;  // uniform over the workgroup
    const int64_t k = (int64_t)e.w + threadIdx.x;
    if (k < e.x || k >= e.y) return;
    eis_bwd_column<T, LM, EX, NT, LAT>(F, Lrt, k, nbp, lowmask, lat, vert, dinv, w, tnb, t);
}

template <typename T, int LM, bool EX, bool WITH_T, bool FUSE, bool DOTS, bool NT, bool SV = true>
__global__ void __launch_bounds__(kBlock) k_eis_fwd_tab(int64_t F, int L, const int4* __restrict__ tab,
                                                        const int32_t* __restrict__ nbp,
                                                        const uint8_t* __restrict__ lowmask, const T* __restrict__ lat,
                                                        const T* __restrict__ vert, const T* __restrict__ dinv,
                                                        const T* __restrict__ E, const T* __restrict__ w,
                                                        const T* __restrict__ t, const T* __restrict__ tnb,
                                                        const T* __restrict__ vnb, T* __restrict__ v,
                                                        T* __restrict__ o, const T* __restrict__ dw,
                                                        double* __restrict__ partial, const double* __restrict__ stop,
                                                        int stop_stride) {
    const int4 e = tab[blockIdx.x];
    if (stop && stop[(int64_t)e.z * stop_stride] != 0.0) return;  // uniform over the workgroup
    const int64_t k = (int64_t)e.w + threadIdx.x;
    double d1 = 0, d2 = 0;
    if (k >= e.x && k < e.y) eis_fwd_column<T, LM, EX, WITH_T, FUSE, DOTS, NT, true, SV>(F, L, k, nbp, lowmask, lat, vert, dinv, E, w,
                                                                            t, tnb, vnb, v, o, dw, d1, d2);
    if constexpr (DOTS) {
        __shared__ double s_red[kBlock / 64];
        const double a = block_sum(d1, s_red);
        const double b = block_sum(d2, s_red);
        if (threadIdx.x == 0) {
            partial[2 * blockIdx.x] = a;
            partial[2 * blockIdx.x + 1] = b;
        }
    }
}

// ---------------------------------------------------------------------------
// Layer-parallel form of the split operator's sweeps.  A workgroup is 64 face
// columns x L layers, one wave per layer (blockDim = (64, L), L <= 16): every
// row's loads and neighbour gathers run in parallel and each wave reads one
// layer, 64 consecutive positions, so every access is coalesced.  Only the
// vertical recurrence is sequential; it is affine in the neighbour value,
//   backward  t_z = a_z - b_z t_{z+1},   a_z = d_z (w_z - U_lat t),  b_z = d_z u_z,
//   forward   v_z = al_z - be_z v_{z-1}, al_z = d_z g_z - t_z,      be_z = d_z l_z,
//             g_z = w_z + E_z t_z - L_lat v  (d = 1/d~, u / l the couplings above / below),
// so wave 0 runs it from (a, b) or (al, be) staged in LDS, and every wave then
// forms its own layer's result o_z = g_z - l_z v_{z-1}.  Short waves and many
// of them: no tail of long-lived column threads, and few registers per lane.
// ---------------------------------------------------------------------------
constexpr int kLpCols = 64;

template <typename T, bool XCD, bool NT>
__global__ void __launch_bounds__(1024) k_eis_bwd_lp(int64_t F, int L, int64_t k0, int64_t k1,
                                                     const int32_t* __restrict__ nbp,
                                                     const uint8_t* __restrict__ lowmask,
                                                     const T* __restrict__ lat, const T* __restrict__ vert,
                                                     const T* __restrict__ dinv, const T* __restrict__ w,
                                                     const T* __restrict__ tnb, T* __restrict__ t,
                                                     const double* __restrict__ stop) {
    if (stop && *stop != 0.0) return;  // a finished BiCGStab run (uniform over the grid)
    extern __shared__ __align__(16) unsigned char lp_smem[];
    T* sa = reinterpret_cast<T*>(lp_smem);
    T* sb = sa + L * kLpCols;
    const int lane = threadIdx.x, z = threadIdx.y;
    const int64_t kb = tile_floor(k0);
    const int64_t k = kb + logical_block<XCD>((k1 - kb + kLpCols - 1) / kLpCols) * kLpCols + lane;
    const bool valid = k >= k0 && k < k1;
    const int64_t N = F * L;
    if (valid) {
        const uint32_t lm = lowmask[k];
        const int64_t q = pos_row(k, z, F, L);
        T acc = 0;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const int32_t n = nbp[f * F + k];
            if (n >= 0 && !((lm >> f) & 1u)) acc += ldc<NT>(&lat[f * N + q]) * tnb[pos_row(n, z, F, L)];
        }
        const T dv = ldc<NT>(&dinv[q]);
        const T va = z < L - 1 ? ldc<NT>(&vert[N + q]) : T(0);
        sa[z * kLpCols + lane] = dv * (w[q] - acc);
        sb[z * kLpCols + lane] = dv * va;
    }
    __syncthreads();
    if (z == 0 && valid) {
        T next = 0;
        for (int zz = L - 1; zz >= 0; --zz) {
            next = sa[zz * kLpCols + lane] - sb[zz * kLpCols + lane] * next;
            t[pos_row(k, zz, F, L)] = next;
        }
    }
}

// FUSE: colour 0 (no lower neighbours), both sweeps in one launch with t kept in
// LDS.  DOTS: partial[2 * logical block] = (dw . o, o . o) over the workgroup (FP64).
template <typename T, bool XCD, bool FUSE, bool DOTS, bool NT>
__global__ void __launch_bounds__(1024) k_eis_fwd_lp(int64_t F, int L, int64_t k0, int64_t k1,
                                                     const int32_t* __restrict__ nbp,
                                                     const uint8_t* __restrict__ lowmask,
                                                     const T* __restrict__ lat, const T* __restrict__ vert,
                                                     const T* __restrict__ dinv, const T* __restrict__ E,
                                                     const T* __restrict__ w, const T* __restrict__ t,
                                                     const T* __restrict__ tnb, const T* __restrict__ vnb,
                                                     T* __restrict__ v, T* __restrict__ o,
                                                     const T* __restrict__ dw, double* __restrict__ partial,
                                                     const double* __restrict__ stop) {
    if (stop && *stop != 0.0) return;  // a finished BiCGStab run (uniform over the grid)
    extern __shared__ __align__(16) unsigned char lp_smem[];
    __shared__ double s_red[2 * 16];
    const int LC = L * kLpCols;
    T* sal = reinterpret_cast<T*>(lp_smem);
    T* sbe = sal + LC;
    T* sv = sbe + LC;
    T* st = sv + LC;  // FUSE only: t, then (a, b) of the backward recurrence
    T* sa = st + LC;
    T* sb = sa + LC;
    const int lane = threadIdx.x, z = threadIdx.y;
    const int64_t kb = tile_floor(k0);
    const int64_t lb = logical_block<XCD>((k1 - kb + kLpCols - 1) / kLpCols);
    const int64_t k = kb + lb * kLpCols + lane;
    const bool valid = k >= k0 && k < k1;
    const int64_t N = F * L;
    const int64_t q = valid ? pos_row(k, z, F, L) : 0;
    const int s = z * kLpCols + lane;
    T wv = 0, dv = 0, lz = 0, ev = 0, g = 0;
    if (valid) {
        const uint32_t lm = lowmask[k];
        wv = w[q];
        dv = ldc<NT>(&dinv[q]);
        lz = z > 0 ? ldc<NT>(&vert[q]) : T(0);
        ev = ldc<NT>(&E[q]);
        if constexpr (FUSE) {
            T acc = 0;
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const int32_t n = nbp[f * F + k];
                if (n >= 0 && !((lm >> f) & 1u)) acc += ldc<NT>(&lat[f * N + q]) * tnb[pos_row(n, z, F, L)];
            }
            const T uz = z < L - 1 ? ldc<NT>(&vert[N + q]) : T(0);
            sa[s] = dv * (wv - acc);
            sb[s] = dv * uz;
        } else {
            T acc = 0;
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const int32_t n = nbp[f * F + k];
                if (n >= 0 && ((lm >> f) & 1u)) acc += ldc<NT>(&lat[f * N + q]) * vnb[pos_row(n, z, F, L)];
            }
            const T tv = t[q];
            g = wv + ev * tv - acc;
            sal[s] = dv * g - tv;
            sbe[s] = dv * lz;
        }
    }
    if constexpr (FUSE) {
        __syncthreads();
        if (z == 0 && valid) {
            T next = 0;
            for (int zz = L - 1; zz >= 0; --zz) {
                next = sa[zz * kLpCols + lane] - sb[zz * kLpCols + lane] * next;
                st[zz * kLpCols + lane] = next;
            }
        }
        __syncthreads();
        if (valid) {
            const T tv = st[s];
            g = wv + ev * tv;
            sal[s] = dv * g - tv;
            sbe[s] = dv * lz;
        }
    }
    __syncthreads();
    if (z == 0 && valid) {
        T prev = 0;
        for (int zz = 0; zz < L; ++zz) {
            prev = sal[zz * kLpCols + lane] - sbe[zz * kLpCols + lane] * prev;
            sv[zz * kLpCols + lane] = prev;
        }
    }
    __syncthreads();
    double d1 = 0, d2 = 0;
    if (valid) {
        const T vz = sv[s];
        const T oz = g - lz * (z > 0 ? sv[s - kLpCols] : T(0));
        v[q] = vz;
        o[q] = oz;
        if (DOTS) {
            d1 = (double)dw[q] * (double)oz;
            d2 = (double)oz * (double)oz;
        }
    }
    if constexpr (DOTS) {
        d1 = wave_sum(d1);
        d2 = wave_sum(d2);
        if (lane == 0) {
            s_red[2 * z] = d1;
            s_red[2 * z + 1] = d2;
        }
        __syncthreads();
        if (z == 0 && lane == 0) {
            double a = 0, b = 0;
            for (int zz = 0; zz < L; ++zz) {
                a += s_red[2 * zz];
                b += s_red[2 * zz + 1];
            }
            partial[2 * lb] = a;
            partial[2 * lb + 1] = b;
        }
    }
}

// y = A x, or r = b - A x with per-block partial sums of r.r, for the operator
// held in position order by the factorisation: diag [N], lat [3N] (neighbour
// j, zero where there is none), vert [2N] (below, above; zero at the ends of
// the column).  One thread per row; blockIdx.y is the layer.
template <typename T, bool RESID>
__global__ void __launch_bounds__(kBlock) k_pos_spmv(int64_t F, int L, const int32_t* __restrict__ nbp,
                                                     const T* __restrict__ diag, const T* __restrict__ lat,
                                                     const T* __restrict__ vert, const T* __restrict__ x,
                                                     const T* __restrict__ b, T* __restrict__ y,
                                                     double* __restrict__ partial) {
    __shared__ double s_red[kBlock / 64];
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    const int64_t N = F * L;
    double rr = 0.0;
    if (k < F) {
        const int64_t qs = pos_stride64(k, F);
        const int64_t q = pos_base64(k, L) + (int64_t)z * qs;
        T acc = diag[q] * x[q];
        if (z > 0) acc = fma(vert[q], x[q - qs], acc);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const int32_t n = nbp[f * F + k];
            const bool on = n >= 0;
            acc = fma(*(on ? &lat[f * N + q] : zero_ptr<T>()), *(on ? &x[pos_row(n, z, F, L)] : zero_ptr<T>()), acc);
        }
        if (z < L - 1) acc = fma(vert[N + q], x[q + qs], acc);
        if (RESID) {
            const T res = b[q] - acc;
            y[q] = res;
            rr = (double)res * (double)res;
        } else {
            y[q] = acc;
        }
    }
    if (RESID) {
        const double sum = block_sum(rr, s_red);
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sum;
    }
}

// y = A x in position order, one thread per face column (L <= LM, fully
// unrolled): the neighbour positions are read once per column and the
// vertical neighbours come from registers.  B is the block size.
// DOTS: the kernel also forms per-block partial sums of w.y and y.y
// (FP64) into partial[2 * block], the block being the physical workgroup
// index (x + y * gridDim.x), so a Krylov step needs no separate dot pass.
template <int B, bool DOTS>
__device__ __forceinline__ void block_dots(double wy, double yy, double* __restrict__ partial) {
    if constexpr (DOTS) {
        __shared__ double s_red[B / 64];
        const double a = block_sum(wy, s_red);
        const double b = block_sum(yy, s_red);
        if (threadIdx.x == 0) {
            const int64_t blk = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
            partial[2 * blk] = a;
            partial[2 * blk + 1] = b;
        }
    }
}

template <typename T, int LM, int B, bool XCD, bool DOTS>
__global__ void __launch_bounds__(B) k_pos_spmv_col(int64_t F, int L, const int32_t* __restrict__ nbp,
                                                    const T* __restrict__ diag, const T* __restrict__ lat,
                                                    const T* __restrict__ vert, const T* __restrict__ x,
                                                    T* __restrict__ y, const T* __restrict__ w,
                                                    double* __restrict__ partial) {
    const int64_t k = logical_block<XCD>((F + B - 1) / B) * B + threadIdx.x;
    double wy = 0, yy = 0;
    if (k < F) {
        const int64_t N = F * L;
        int32_t np_[3];
        int64_t nq[3], ns[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            np_[f] = nbp[f * F + k];
            nq[f] = pos_base64(max(np_[f], 0), L);
            ns[f] = pos_stride64(max(np_[f], 0), F);
        }
        const int64_t qb = pos_base64(k, L), qs = pos_stride64(k, F);
        T xv[LM];
#pragma unroll
        for (int z = 0; z < LM; ++z)
            if (z < L) xv[z] = x[qb + (int64_t)z * qs];
#pragma unroll
        for (int z = 0; z < LM; ++z)
            if (z < L) {
                const int64_t q = qb + (int64_t)z * qs;
                T acc = diag[q] * xv[z];
                if (z > 0) acc = fma(vert[q], xv[z > 0 ? z - 1 : 0], acc);
#pragma unroll
                for (int f = 0; f < 3; ++f)
                {
                    const bool on = np_[f] >= 0;
                    acc = fma(*(on ? &lat[f * N + q] : zero_ptr<T>()), *(on ? &x[nq[f] + (int64_t)z * ns[f]] : zero_ptr<T>()),
                              acc);
                }
                if (z < L - 1) acc = fma(vert[N + q], xv[z + 1 < LM ? z + 1 : z], acc);
                y[q] = acc;
                if (DOTS) {
                    wy = fma((double)w[q], (double)acc, wy);
                    yy = fma((double)acc, (double)acc, yy);
                }
            }
    }
    block_dots<B, DOTS>(wy, yy, partial);
}

// y = A x in position order, one thread per row, block size B (grid: F/B x L)
template <typename T, int B, bool XCD, bool DOTS>
__global__ void __launch_bounds__(B) k_pos_spmv_row(int64_t F, int L, const int32_t* __restrict__ nbp,
                                                    const T* __restrict__ diag, const T* __restrict__ lat,
                                                    const T* __restrict__ vert, const T* __restrict__ x,
                                                    T* __restrict__ y, const T* __restrict__ w,
                                                    double* __restrict__ partial) {
    const int64_t k = logical_block<XCD>((F + B - 1) / B) * B + threadIdx.x;
    double wy = 0, yy = 0;
    if (k < F) {
        const int z = blockIdx.y;
        const int64_t N = F * L;
        const int64_t qs = pos_stride64(k, F);
        const int64_t q = pos_base64(k, L) + (int64_t)z * qs;
        T acc = diag[q] * x[q];
        if (z > 0) acc = fma(vert[q], x[q - qs], acc);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const int32_t n = nbp[f * F + k];
            const bool on = n >= 0;
            acc = fma(*(on ? &lat[f * N + q] : zero_ptr<T>()), *(on ? &x[pos_row(n, z, F, L)] : zero_ptr<T>()), acc);
        }
        if (z < L - 1) acc = fma(vert[N + q], x[q + qs], acc);
        y[q] = acc;
        if (DOTS) {
            wy = (double)w[q] * (double)acc;
            yy = (double)acc * (double)acc;
        }
    }
    block_dots<B, DOTS>(wy, yy, partial);
}

// reference order (z*F + face) <-> position order (tiled rows), with a type change;
// blockIdx.y is the layer (of L = gridDim.y)
template <typename Ti, typename To>
__global__ void __launch_bounds__(kBlock) k_to_pos(int64_t F, const int32_t* __restrict__ faces,
                                                   const Ti* __restrict__ in, To* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= F) return;
    const int z = blockIdx.y;
    out[pos_row(k, z, F, gridDim.y)] = (To)in[(int64_t)z * F + faces[k]];
}
template <typename Ti, typename To>
__global__ void __launch_bounds__(kBlock) k_from_pos(int64_t F, const int32_t* __restrict__ faces,
                                                     const Ti* __restrict__ in, To* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= F) return;
    const int z = blockIdx.y;
    out[(int64_t)z * F + faces[k]] = (To)in[pos_row(k, z, F, gridDim.y)];
}

__global__ void __launch_bounds__(kBlock) k_narrow(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

}  // namespace

// Colouring of the face graph (every face has at most 3 neighbours).  Greedy in
// index order gives at most 4 colours; faces left with colour 3 are then
// repaired in index order (Brooks: graphs of maximum degree 3 other than K4 are
// 3-colourable).  repair(v): take a colour below 3 that no neighbour holds, else
// for the colour pairs (ca, cb) = (0, 1), (0, 2), (1, 2) in turn, with a / b the
// neighbours of colours ca / cb, collect the connected set of faces reachable
// from a through colours {ca, cb} (at most kKempeCap faces); if b is not in it,
// swap ca and cb on the set (a Kempe swap) and give v colour ca.  A face that
// repair cannot fix trades colours with a neighbour (v takes the neighbour's
// colour c, for c = 0, 1, 2 in turn) when repair then fixes that neighbour.  A
// fourth colour costs two nearly empty launches per preconditioner apply.
constexpr int64_t kKempeCap = 4096;

static std::vector<int32_t> color_faces(int64_t F, const int32_t* nb) {
    std::vector<int32_t> color(F, -1);
    for (int64_t i = 0; i < F; ++i) {
        unsigned used = 0;
        for (int f = 0; f < 3; ++f) {
            const int32_t n = nb[3 * i + f];
            if (n >= 0 && color[n] >= 0) used |= 1u << color[n];
        }
        int c = 0;
        while (used & (1u << c)) ++c;
        color[i] = c;
    }
    std::vector<int64_t> stamp(F, -1), chain;
    int64_t visit = 0;
    auto repair = [&](int64_t v) -> bool {
        unsigned used = 0;
        for (int f = 0; f < 3; ++f)
            if (nb[3 * v + f] >= 0) used |= 1u << color[nb[3 * v + f]];
        for (int c = 0; c < 3; ++c)
            if (!(used & (1u << c))) {
                color[v] = c;
                return true;
            }
        int32_t by_color[3] = {-1, -1, -1};
        for (int f = 0; f < 3; ++f) by_color[color[nb[3 * v + f]]] = nb[3 * v + f];
        static const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (int pi = 0; pi < 3; ++pi) {
            const int32_t ca = pairs[pi][0], cb = pairs[pi][1];
            const int32_t a = by_color[ca], b = by_color[cb];
            ++visit;
            chain.clear();
            chain.push_back(a);
            stamp[a] = visit;
            bool hit_b = false;
            for (size_t h = 0; h < chain.size() && !hit_b && (int64_t)chain.size() <= kKempeCap; ++h) {
                const int32_t u = (int32_t)chain[h];
                for (int f = 0; f < 3; ++f) {
                    const int32_t w = nb[3 * (int64_t)u + f];
                    if (w < 0 || stamp[w] == visit) continue;
                    if (color[w] != ca && color[w] != cb) continue;
                    if (w == b) hit_b = true;
                    stamp[w] = visit;
                    chain.push_back(w);
                }
            }
            if (hit_b || (int64_t)chain.size() > kKempeCap) continue;
            for (int64_t u : chain) color[u] = color[u] == ca ? cb : ca;
            color[v] = ca;
            return true;
        }
        return false;
    };
    for (int64_t v = 0; v < F; ++v)
        if (color[v] == 3 && !repair(v)) {
            int32_t by_color[3] = {-1, -1, -1};
            for (int f = 0; f < 3; ++f) by_color[color[nb[3 * v + f]]] = nb[3 * v + f];
            for (int c = 0; c < 3; ++c) {
                const int32_t a = by_color[c];
                color[v] = c;
                color[a] = 3;
                if (repair(a)) break;
                color[a] = c;
                color[v] = 3;
            }
        }
    return color;
}

int face_colors(int64_t F, const int32_t* nb, int32_t* color) {
    const std::vector<int32_t> c = color_faces(F, nb);
    int nc = 0;
    for (int64_t i = 0; i < F; ++i) {
        color[i] = c[i];
        nc = std::max(nc, c[i] + 1);
    }
    return nc;
}

// Position order of the multicolour D-ILU.  Colours are laid out one after the other.
// Within a colour, faces are ordered by connected component of the face graph (the members
// of a batch union stay contiguous) and face index, and then -- for the middle colours, which
// have both lower- and higher-colour neighbours -- regrouped inside blocks of 1024 faces by
// how many of their neighbours carry a lower colour.  Each face's three neighbour slots hold its lower-colour neighbours first,
// then the higher-colour ones, then the absent ones.  The forward sweep reads the lateral
// couplings of the lower slots and the backward sweep those of the higher slots; with the
// faces of one count together, the cache lines of a slot's coupling array hold only lower or
// only higher couplings over each run, so each sweep fetches its own couplings and not the
// other sweep's.  Faces of one colour are not coupled to each other, so their order does not
// change the factorisation.
void build_color_order(ColorOrder& co, int64_t F, const int32_t* nb, hipStream_t s) {
    std::vector<int32_t> color = color_faces(F, nb);
    int nc = 0;
    for (int64_t i = 0; i < F; ++i) nc = std::max(nc, color[i] + 1);
    co.ncolor = nc;
    co.off.assign(nc + 1, 0);
    for (int64_t i = 0; i < F; ++i) co.off[color[i] + 1]++;
    for (int c = 0; c < nc; ++c) co.off[c + 1] += co.off[c];
    // connected components, numbered in the order of their smallest face index
    std::vector<int32_t> comp(F, -1), stack;
    int32_t ncomp = 0;
    for (int64_t i0 = 0; i0 < F; ++i0) {
        if (comp[i0] >= 0) continue;
        comp[i0] = ncomp;
        stack.assign(1, (int32_t)i0);
        while (!stack.empty()) {
            const int32_t u = stack.back();
            stack.pop_back();
            for (int f = 0; f < 3; ++f) {
                const int32_t w = nb[3 * (int64_t)u + f];
                if (w >= 0 && comp[w] < 0) {
                    comp[w] = ncomp;
                    stack.push_back(w);
                }
            }
        }
        ++ncomp;
    }
    std::vector<uint8_t> nlo(F, 0);
    for (int64_t i = 0; i < F; ++i)
        for (int f = 0; f < 3; ++f) {
            const int32_t n = nb[3 * i + f];
            if (n >= 0 && color[n] < color[i]) nlo[i]++;
        }
    // the regrouping by lower-neighbour count stays inside blocks of kGroupBlock consecutive
    // faces of a colour: a block covers a compact patch of the mesh, so the neighbours its
    // faces gather stay within one patch as well
    constexpr int64_t kGroupBlock = 1024;
    std::vector<int32_t> faces(F), pos(F), blk(F);
    for (int64_t i = 0; i < F; ++i) faces[i] = (int32_t)i;
    std::stable_sort(faces.begin(), faces.end(), [&](int32_t a, int32_t b) {
        if (color[a] != color[b]) return color[a] < color[b];
        if (comp[a] != comp[b]) return comp[a] < comp[b];
        return a < b;
    });
    for (int64_t k = 0; k < F; ++k) blk[faces[k]] = (int32_t)((k - co.off[color[faces[k]]]) / kGroupBlock);
    std::stable_sort(faces.begin(), faces.end(), [&](int32_t a, int32_t b) {
        if (color[a] != color[b]) return color[a] < color[b];
        if (comp[a] != comp[b]) return comp[a] < comp[b];
        if (blk[a] != blk[b]) return blk[a] < blk[b];
        const bool mid = color[a] > 0 && color[a] < nc - 1;
        if (mid && nlo[a] != nlo[b]) return nlo[a] < nlo[b];
        return a < b;
    });
    for (int64_t k = 0; k < F; ++k) pos[faces[k]] = (int32_t)k;
    std::vector<int32_t> nbp(3 * F, -1);
    std::vector<uint8_t> rev(3 * F, 0), low(F, 0), edge(F, 0);
    for (int64_t k = 0; k < F; ++k) {
        const int64_t i = faces[k];
        // slot order: lower-colour neighbours, higher-colour ones, absent ones (edge order within)
        int e[3], ne = 0;
        for (int pass = 0; pass < 3; ++pass)
            for (int f = 0; f < 3; ++f) {
                const int32_t n = nb[3 * i + f];
                const int cls = n < 0 ? 2 : (color[n] < color[i] ? 0 : 1);
                if (cls == pass) e[ne++] = f;
            }
        for (int sl = 0; sl < 3; ++sl) {
            const int f = e[sl];
            edge[k] |= (uint8_t)(f << (2 * sl));
            const int32_t n = nb[3 * i + f];
            if (n < 0) continue;
            nbp[sl * F + k] = pos[n];
            for (int g = 0; g < 3; ++g)
                if (nb[3 * (int64_t)n + g] == (int32_t)i) rev[sl * F + k] = (uint8_t)g;
            if (color[n] < color[i]) low[k] |= (uint8_t)(1u << sl);
        }
    }
    co.faces.alloc(F);
    co.nbp.alloc(3 * F);
    co.rev.alloc(3 * F);
    co.lowmask.alloc(F);
    co.edge.alloc(F);
    PB_HIP(hipMemcpyAsync(co.faces.p, faces.data(), F * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(co.nbp.p, nbp.data(), 3 * F * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(co.rev.p, rev.data(), 3 * F, hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(co.lowmask.p, low.data(), F, hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(co.edge.p, edge.data(), F, hipMemcpyHostToDevice, s));
    PB_HIP(hipStreamSynchronize(s));
    co.h_color = std::move(color);
    co.h_faces = std::move(faces);
}

template <typename T>
void dilu_factor(hipStream_t s, const ColorOrder& co, int64_t F, int L, const int32_t* rp, const T* val,
                 const uint16_t* slots, DiluFactors<T>& d) {
    for (int c = 0; c < co.ncolor; ++c) {
        const int64_t k0 = co.off[c], k1 = co.off[c + 1];
        if (k1 == k0) continue;
        hipLaunchKernelGGL((k_dilu_factor<T>), dim3(grid_for(k1 - k0)), dim3(kBlock), 0, s, F, L, k0, k1, co.faces.p,
                           co.nbp.p, co.rev.p, co.lowmask.p, co.edge.p, rp, val, slots, d.lat.p, d.vert.p, d.dinv.p,
                           d.diag.p, d.E.p);
    }
}

namespace {
template <typename T, bool POS, int LM, bool XCD>
void dilu_apply_t(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* r, T* z) {
    const int top = co.ncolor - 1;
    for (int c = 0; c <= top; ++c) {
        const int64_t k0 = co.off[c], k1 = co.off[c + 1];
        if (k1 == k0) continue;
        const dim3 g((unsigned)launch_blocks<XCD>(col_blocks(k0, k1)));
        if (c == top)
            hipLaunchKernelGGL((k_dilu_fwd<T, POS, LM, true, XCD>), g, dim3(kBlock), 0, s, F, L, k0, k1, co.faces.p,
                               co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, r, d.y.p, d.y.p, z);
        else
            hipLaunchKernelGGL((k_dilu_fwd<T, POS, LM, false, XCD>), g, dim3(kBlock), 0, s, F, L, k0, k1, co.faces.p,
                               co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, r, d.y.p, d.y.p, z);
    }
    // reference order: higher-colour results are read back from y (updated in place)
    const T* done = POS ? z : d.y.p;
    for (int c = top - 1; c >= 0; --c) {
        const int64_t k0 = co.off[c], k1 = co.off[c + 1];
        if (k1 == k0) continue;
        const dim3 g((unsigned)launch_blocks<XCD>(col_blocks(k0, k1)));
        hipLaunchKernelGGL((k_dilu_bwd<T, POS, LM, XCD>), g, dim3(kBlock), 0, s, F, L, k0, k1, co.faces.p, co.nbp.p,
                           co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, d.y.p, done, z);
    }
}
template <typename T, bool POS, bool XCD>
void dilu_apply_l(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* r, T* z) {
    if (L <= 1)
        dilu_apply_t<T, POS, 1, XCD>(s, co, F, L, d, r, z);
    else if (L <= 5)
        dilu_apply_t<T, POS, 5, XCD>(s, co, F, L, d, r, z);
    else if (L <= 10)
        dilu_apply_t<T, POS, 10, XCD>(s, co, F, L, d, r, z);
    else if (L <= 16)
        dilu_apply_t<T, POS, 16, XCD>(s, co, F, L, d, r, z);
    else
        dilu_apply_t<T, POS, 0, XCD>(s, co, F, L, d, r, z);
}
}  // namespace

template <typename T>
void dilu_apply(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* r, T* z, bool pos,
                bool xcd) {
    if (pos && xcd)
        dilu_apply_l<T, true, true>(s, co, F, L, d, r, z);
    else if (pos)
        dilu_apply_l<T, true, false>(s, co, F, L, d, r, z);
    else if (xcd)
        dilu_apply_l<T, false, true>(s, co, F, L, d, r, z);
    else
        dilu_apply_l<T, false, false>(s, co, F, L, d, r, z);
}

template <typename T>
bool tune_dilu_xcd(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* r, T* z,
                   float* ms_out) {
    hipEvent_t e0, e1;
    PB_HIP(hipEventCreate(&e0));
    PB_HIP(hipEventCreate(&e1));
    float ms[2] = {0, 0};
    for (int v = 0; v < 2; ++v) {
        dilu_apply<T>(s, co, F, L, d, r, z, true, v == 1);
        PB_HIP(hipEventRecord(e0, s));
        for (int w = 0; w < 3; ++w) dilu_apply<T>(s, co, F, L, d, r, z, true, v == 1);
        PB_HIP(hipEventRecord(e1, s));
        PB_HIP(hipEventSynchronize(e1));
        PB_HIP(hipEventElapsedTime(&ms[v], e0, e1));
        ms[v] /= 3;
        if (ms_out) ms_out[v] = ms[v];
    }
    PB_HIP(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms[1] < ms[0];
}

namespace {
// one forward launch of colour c; with dw the block partials of (dw . o, o . o) go to
// partial[2 * (nb + block)] and nb advances by the launch's block count
// SV = false for the top colour: its v is not stored, no higher colour gathers it
template <typename T, int LM, bool EX, bool XCD, bool NT, bool WITH_T, bool FUSE, bool SV>
void launch_fwd(hipStream_t s, dim3 g, int64_t F, int L, const ColorOrder& co, int c, DiluFactors<T>& d, const T* w,
                T* o, const T* dw, double* partial, int64_t& nb, const double* stop) {
    if (dw) {
        hipLaunchKernelGGL((k_eis_fwd<T, LM, EX, XCD, WITH_T, FUSE, true, NT, SV>), g, dim3(kBlock), 0, s, F, L,
                           co.off[c], co.off[c + 1], co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, d.E.p, w,
                           d.t.p, d.t.p, d.y.p, d.y.p, o, dw, partial + 2 * nb, stop);
        nb += g.x;
    } else {
        hipLaunchKernelGGL((k_eis_fwd<T, LM, EX, XCD, WITH_T, FUSE, false, NT, SV>), g, dim3(kBlock), 0, s, F, L,
                           co.off[c], co.off[c + 1], co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, d.E.p, w,
                           d.t.p, d.t.p, d.y.p, d.y.p, o, dw, partial, stop);
    }
}

template <typename T, int LM, bool EX, bool XCD, bool NT>
int64_t split_apply_t(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* w, T* o,
                      const T* dw, double* partial, const double* stop) {
    int64_t nb = 0;
    const int top = co.ncolor - 1;
    // grids rounded to a multiple of the XCD count with or without the XCD mapping: the dot
    // partials then have the same layout in both (the extra workgroups add zeros)
    auto blocks = [&](int c) { return dim3((unsigned)launch_blocks<true>(col_blocks(co.off[c], co.off[c + 1]))); };
    // backward sweep, colours top .. 1: t (the top colour has no upper lateral neighbours)
    for (int c = top; c >= 1; --c) {
        if (co.off[c + 1] == co.off[c]) continue;
        if (c == top)
            hipLaunchKernelGGL((k_eis_bwd<T, LM, EX, XCD, NT, false>), blocks(c), dim3(kBlock), 0, s, F, L, co.off[c],
                               co.off[c + 1], co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, w, d.t.p, d.t.p, stop);
        else
            hipLaunchKernelGGL((k_eis_bwd<T, LM, EX, XCD, NT>), blocks(c), dim3(kBlock), 0, s, F, L, co.off[c],
                               co.off[c + 1], co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, w, d.t.p, d.t.p, stop);
    }
    // colour 0: backward and forward sweeps in one pass
    if (co.off[1] > co.off[0])
        launch_fwd<T, LM, EX, XCD, NT, true, true, true>(s, blocks(0), F, L, co, 0, d, w, o, dw, partial, nb, stop);
    // forward sweep, colours 1 .. top
    for (int c = 1; c <= top; ++c) {
        if (co.off[c + 1] == co.off[c]) continue;
        if (c == top)
            launch_fwd<T, LM, EX, XCD, NT, true, false, false>(s, blocks(c), F, L, co, c, d, w, o, dw, partial, nb,
                                                               stop);
        else
            launch_fwd<T, LM, EX, XCD, NT, true, false, true>(s, blocks(c), F, L, co, c, d, w, o, dw, partial, nb,
                                                              stop);
    }
    return nb;
}
template <typename T, int LM, bool XCD, bool NT>
void split_lsolve_t(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* b, T* bhat) {
    for (int c = 0; c < co.ncolor; ++c) {
        if (co.off[c + 1] == co.off[c]) continue;
        const dim3 g((unsigned)launch_blocks<XCD>(col_blocks(co.off[c], co.off[c + 1])));
        hipLaunchKernelGGL((k_eis_fwd<T, LM, false, XCD, false, false, false, NT>), g, dim3(kBlock), 0, s, F, L, co.off[c],
                           co.off[c + 1], co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, d.E.p, b, d.t.p,
                           d.t.p, d.y.p, d.y.p, bhat, (const T*)nullptr, (double*)nullptr, (const double*)nullptr);
    }
}
template <typename T, int LM, bool XCD, bool NT>
void split_rsolve_t(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* xhat, T* x) {
    for (int c = co.ncolor - 1; c >= 0; --c) {
        if (co.off[c + 1] == co.off[c]) continue;
        const dim3 g((unsigned)launch_blocks<XCD>(col_blocks(co.off[c], co.off[c + 1])));
        hipLaunchKernelGGL((k_eis_bwd<T, LM, false, XCD, NT>), g, dim3(kBlock), 0, s, F, L, co.off[c], co.off[c + 1],
                           co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, xhat, x, x, (const double*)nullptr);
    }
}
// which: 0 apply, 1 left transform, 2 right transform
template <typename T, bool XCD, bool NT>
int64_t split_dispatch(int which, hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d,
                       const T* in, T* out, const T* dw, double* partial, const double* stop) {
    int64_t nb = 0;
#define PB_SPLIT(LMV)                                                           \
    {                                                                           \
        if (which == 0)                                                         \
            nb = L == LMV ? split_apply_t<T, LMV, true, XCD, NT>(s, co, F, L, d, in, out, dw, partial, stop) \
                          : split_apply_t<T, LMV, false, XCD, NT>(s, co, F, L, d, in, out, dw, partial, stop); \
        else if (which == 1)                                                    \
            split_lsolve_t<T, LMV, XCD, NT>(s, co, F, L, d, in, out);           \
        else                                                                    \
            split_rsolve_t<T, LMV, XCD, NT>(s, co, F, L, d, in, out);           \
    }
    if (L <= 1)
        PB_SPLIT(1)
    else if (L <= 5)
        PB_SPLIT(5)
    else if (L <= 10)
        PB_SPLIT(10)
    else if (L <= 16)
        PB_SPLIT(16)
    else
        throw Error(PBSM3D_EINVAL, "the split D-ILU operator needs nLayer <= 16");
#undef PB_SPLIT
    return nb;
}
}  // namespace

namespace {
// the split operator with the layer-parallel sweeps (L <= 16)
template <typename T, bool XCD, bool NT>
int64_t split_apply_lp(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* w, T* o,
                       const T* dw, double* partial, const double* stop) {
    if (L < 1 || L > 16) throw Error(PBSM3D_EINVAL, "the split D-ILU operator needs nLayer <= 16");
    int64_t nb = 0;
    const int top = co.ncolor - 1;
    const dim3 blk(kLpCols, (unsigned)L);
    const size_t lc = (size_t)L * kLpCols * sizeof(T);
    auto blocks = [&](int c) {  // rounded as in split_apply_t
        return dim3((unsigned)launch_blocks<true>(col_blocks(co.off[c], co.off[c + 1], kLpCols)));
    };
    for (int c = top; c >= 1; --c) {
        if (co.off[c + 1] == co.off[c]) continue;
        hipLaunchKernelGGL((k_eis_bwd_lp<T, XCD, NT>), blocks(c), blk, 2 * lc, s, F, L, co.off[c], co.off[c + 1],
                           co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, w, d.t.p, d.t.p, stop);
    }
    for (int c = 0; c <= top; ++c) {
        if (co.off[c + 1] == co.off[c]) continue;
        const dim3 g = blocks(c);
        double* part = dw ? partial + 2 * nb : nullptr;
#define PB_LP_FWD(FUSE, DOTS)                                                                                      \
    hipLaunchKernelGGL((k_eis_fwd_lp<T, XCD, FUSE, DOTS, NT>), g, blk, (FUSE ? 6 : 3) * lc, s, F, L, co.off[c],   \
                       co.off[c + 1], co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, d.E.p, w, d.t.p, d.t.p, \
                       d.y.p, d.y.p, o, dw, part, stop)
        if (c == 0) {
            if (dw)
                PB_LP_FWD(true, true);
            else
                PB_LP_FWD(true, false);
        } else {
            if (dw)
                PB_LP_FWD(false, true);
            else
                PB_LP_FWD(false, false);
        }
#undef PB_LP_FWD
        if (dw) nb += g.x;
    }
    return nb;
}

template <typename T>
int64_t split_mode_dispatch(int which, hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d,
                            const T* in, T* out, const T* dw, double* partial, int mode, const double* stop) {
    if (which == 0 && (mode & 4)) {  // layer-parallel sweeps (the transforms keep the column kernels)
        switch (mode & 3) {
            case 0: return split_apply_lp<T, false, false>(s, co, F, L, d, in, out, dw, partial, stop);
            case 1: return split_apply_lp<T, true, false>(s, co, F, L, d, in, out, dw, partial, stop);
            case 2: return split_apply_lp<T, false, true>(s, co, F, L, d, in, out, dw, partial, stop);
            default: return split_apply_lp<T, true, true>(s, co, F, L, d, in, out, dw, partial, stop);
        }
    }
    switch (mode & 3) {
        case 0: return split_dispatch<T, false, false>(which, s, co, F, L, d, in, out, dw, partial, stop);
        case 1: return split_dispatch<T, true, false>(which, s, co, F, L, d, in, out, dw, partial, stop);
        case 2: return split_dispatch<T, false, true>(which, s, co, F, L, d, in, out, dw, partial, stop);
        default: return split_dispatch<T, true, true>(which, s, co, F, L, d, in, out, dw, partial, stop);
    }
}
}  // namespace

// the column sweeps index positions in 32 bits, pad slots included
inline void check_split_size(int64_t F, int L) {
    if ((uint64_t)3 * (uint64_t)F * (uint64_t)L + 1 >= (1ull << 32))
        throw Error(PBSM3D_EINVAL, "the split D-ILU operator needs 3 * faces * nLayer < 2^32");
}

template <typename T>
void split_op(int which, hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* in, T* out,
              int mode) {
    check_split_size(F, L);
    split_mode_dispatch<T>(which, s, co, F, L, d, in, out, nullptr, nullptr, mode, nullptr);
}

template <typename T>
int64_t split_apply_dots(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* in, T* out,
                         const T* dw, double* partial, int mode, const double* stop) {
    check_split_size(F, L);
    return split_mode_dispatch<T>(0, s, co, F, L, d, in, out, dw, partial, mode, stop);
}

namespace {
template <typename T, int LM, bool EX>
void split_apply_batch_t(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d,
                         const SweepTables& st, const T* w, T* o, const T* dw, double* partial, const double* stop,
                         int stop_stride) {
    const int top = co.ncolor - 1;
    for (int c = top; c >= 1; --c) {
        if (st.nblk[c] == 0) continue;
        if (c == top)
            hipLaunchKernelGGL((k_eis_bwd_tab<T, LM, EX, false, false>), dim3((unsigned)st.nblk[c]), dim3(kBlock), 0, s,
                               F, L, st.tab[c].p, co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, w, d.t.p, d.t.p,
                               stop, stop_stride);
        else
            hipLaunchKernelGGL((k_eis_bwd_tab<T, LM, EX, false>), dim3((unsigned)st.nblk[c]), dim3(kBlock), 0, s, F, L,
                               st.tab[c].p, co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, w, d.t.p, d.t.p, stop,
                               stop_stride);
    }
    for (int c = 0; c <= top; ++c) {
        if (st.nblk[c] == 0) continue;
        const dim3 g((unsigned)st.nblk[c]);
        double* pc = dw ? partial + 2 * st.part_off[c] : nullptr;
#define PB_FWD_TAB(FUSE, DOTS, SV)                                                                                 \
    hipLaunchKernelGGL((k_eis_fwd_tab<T, LM, EX, true, FUSE, DOTS, false, SV>), g, dim3(kBlock), 0, s, F, L,         \
                       st.tab[c].p, co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, d.E.p, w, d.t.p, d.t.p, d.y.p, \
                       d.y.p, o, dw, pc, stop, stop_stride)
        // the top colour does not store v (no higher colour gathers it)
        if (c == 0) {
            if (dw) PB_FWD_TAB(true, true, true); else PB_FWD_TAB(true, false, true);
        } else if (c == top) {
            if (dw) PB_FWD_TAB(false, true, false); else PB_FWD_TAB(false, false, false);
        } else {
            if (dw) PB_FWD_TAB(false, true, true); else PB_FWD_TAB(false, false, true);
        }
#undef PB_FWD_TAB
    }
}
}  // namespace

template <typename T>
void split_apply_batch(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const SweepTables& st,
                       const T* in, T* out, const T* dw, double* partial, const double* stop, int stop_stride) {
    if (co.ncolor > kMaxColors) throw Error(PBSM3D_EINVAL, "too many colours for the batch tables");
    check_split_size(F, L);
#define PB_BATCH(LMV)                                                                                   \
    {                                                                                                   \
        if (L == LMV)                                                                                   \
            split_apply_batch_t<T, LMV, true>(s, co, F, L, d, st, in, out, dw, partial, stop, stop_stride); \
        else                                                                                            \
            split_apply_batch_t<T, LMV, false>(s, co, F, L, d, st, in, out, dw, partial, stop, stop_stride); \
    }
    if (L <= 1)
        PB_BATCH(1)
    else if (L <= 5)
        PB_BATCH(5)
    else if (L <= 10)
        PB_BATCH(10)
    else if (L <= 16)
        PB_BATCH(16)
    else
        throw Error(PBSM3D_EINVAL, "the split D-ILU operator needs nLayer <= 16");
#undef PB_BATCH
    PB_HIP(hipGetLastError());
}

template <typename T>
int tune_split_mode(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* w, T* o,
                    float* ms_out) {
    hipEvent_t e0, e1;
    PB_HIP(hipEventCreate(&e0));
    PB_HIP(hipEventCreate(&e1));
    int best = 0;
    float best_ms = 0;
    for (int m = 0; m < kSplitModes; ++m) {
        float ms = 0;
        split_op<T>(0, s, co, F, L, d, w, o, m);
        PB_HIP(hipEventRecord(e0, s));
        for (int k = 0; k < 3; ++k) split_op<T>(0, s, co, F, L, d, w, o, m);
        PB_HIP(hipEventRecord(e1, s));
        PB_HIP(hipEventSynchronize(e1));
        PB_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= 3;
        if (ms_out) ms_out[m] = ms;
        if (m == 0 || ms < best_ms) {
            best = m;
            best_ms = ms;
        }
    }
    PB_HIP(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

void dilu_narrow(hipStream_t s, int64_t N, const DiluFactors<double>& a, DiluFactors<float>& b) {
    hipLaunchKernelGGL(k_narrow, dim3(grid_for(3 * N)), dim3(kBlock), 0, s, 3 * N, a.lat.p, b.lat.p);
    hipLaunchKernelGGL(k_narrow, dim3(grid_for(2 * N)), dim3(kBlock), 0, s, 2 * N, a.vert.p, b.vert.p);
    hipLaunchKernelGGL(k_narrow, dim3(grid_for(N)), dim3(kBlock), 0, s, N, a.dinv.p, b.dinv.p);
    hipLaunchKernelGGL(k_narrow, dim3(grid_for(N)), dim3(kBlock), 0, s, N, a.diag.p, b.diag.p);
    hipLaunchKernelGGL(k_narrow, dim3(grid_for(N)), dim3(kBlock), 0, s, N, a.E.p, b.E.p);
}

namespace {
// the launch shapes the tuner chooses from: shape v % 5 (0 / 4 one thread per
// row with blocks of 256 / 128; 1 / 2 / 3 one thread per face column with
// blocks of 256 / 128 / 64), with the XCD workgroup mapping when v >= 5
template <typename T, int B, bool XCD, bool DOTS>
int64_t row_launch(hipStream_t s, int64_t F, int L, const ColorOrder& co, const DiluFactors<T>& d, const T* x, T* y,
                   const T* w, double* partial) {
    const dim3 g((unsigned)launch_blocks<XCD>((F + B - 1) / B), (unsigned)L);
    hipLaunchKernelGGL((k_pos_spmv_row<T, B, XCD, DOTS>), g, dim3(B), 0, s, F, L, co.nbp.p, d.diag.p, d.lat.p,
                       d.vert.p, x, y, w, partial);
    return (int64_t)g.x * g.y;
}
template <typename T, int LM, int B, bool XCD, bool DOTS>
int64_t col_launch(hipStream_t s, int64_t F, int L, const ColorOrder& co, const DiluFactors<T>& d, const T* x, T* y,
                   const T* w, double* partial) {
    const dim3 g((unsigned)launch_blocks<XCD>((F + B - 1) / B));
    hipLaunchKernelGGL((k_pos_spmv_col<T, LM, B, XCD, DOTS>), g, dim3(B), 0, s, F, L, co.nbp.p, d.diag.p, d.lat.p,
                       d.vert.p, x, y, w, partial);
    return (int64_t)g.x;
}
template <typename T, int LM, bool XCD, bool DOTS>
int64_t col_shapes(hipStream_t s, int shape, int64_t F, int L, const ColorOrder& co, const DiluFactors<T>& d,
                   const T* x, T* y, const T* w, double* partial) {
    if (shape == 1) return col_launch<T, LM, 256, XCD, DOTS>(s, F, L, co, d, x, y, w, partial);
    if (shape == 2) return col_launch<T, LM, 128, XCD, DOTS>(s, F, L, co, d, x, y, w, partial);
    return col_launch<T, LM, 64, XCD, DOTS>(s, F, L, co, d, x, y, w, partial);
}
template <typename T, bool XCD, bool DOTS>
int64_t shapes(hipStream_t s, int shape, int64_t F, int L, const ColorOrder& co, const DiluFactors<T>& d, const T* x,
               T* y, const T* w, double* partial) {
    if (shape >= 1 && shape <= 3 && L <= 16) {
        if (L <= 5) return col_shapes<T, 5, XCD, DOTS>(s, shape, F, L, co, d, x, y, w, partial);
        if (L <= 10) return col_shapes<T, 10, XCD, DOTS>(s, shape, F, L, co, d, x, y, w, partial);
        return col_shapes<T, 16, XCD, DOTS>(s, shape, F, L, co, d, x, y, w, partial);
    }
    if (shape == 4) return row_launch<T, 128, XCD, DOTS>(s, F, L, co, d, x, y, w, partial);
    return row_launch<T, 256, XCD, DOTS>(s, F, L, co, d, x, y, w, partial);
}
bool shape_allowed(int v, int L) {
    const int shape = v % 5;
    return !(shape >= 1 && shape <= 3 && L > 16);
}
}  // namespace

template <typename T>
void pos_spmv_variant(hipStream_t s, int v, const ColorOrder& co, int64_t F, int L, const DiluFactors<T>& d,
                      const T* x, T* y) {
    if (v >= 5)
        shapes<T, true, false>(s, v - 5, F, L, co, d, x, y, nullptr, nullptr);
    else
        shapes<T, false, false>(s, v, F, L, co, d, x, y, nullptr, nullptr);
}

template <typename T>
int64_t pos_spmv_dots(hipStream_t s, int v, const ColorOrder& co, int64_t F, int L, const DiluFactors<T>& d,
                      const T* x, T* y, const T* w, double* partial) {
    if (v >= 5) return shapes<T, true, true>(s, v - 5, F, L, co, d, x, y, w, partial);
    return shapes<T, false, true>(s, v, F, L, co, d, x, y, w, partial);
}

template <typename T>
int tune_pos_spmv(hipStream_t s, const ColorOrder& co, int64_t F, int L, const DiluFactors<T>& d, const T* x, T* y,
                  float* ms_out) {
    hipEvent_t e0, e1;
    PB_HIP(hipEventCreate(&e0));
    PB_HIP(hipEventCreate(&e1));
    int best = 0;
    float best_ms = 1e30f;
    for (int v = 0; v < kPosSpmvVariants; ++v) {
        if (!shape_allowed(v, L)) continue;
        for (int w = 0; w < 2; ++w) pos_spmv_variant<T>(s, v, co, F, L, d, x, y);
        PB_HIP(hipEventRecord(e0, s));
        for (int w = 0; w < 5; ++w) pos_spmv_variant<T>(s, v, co, F, L, d, x, y);
        PB_HIP(hipEventRecord(e1, s));
        PB_HIP(hipEventSynchronize(e1));
        float ms = 0;
        PB_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= 5;
        if (ms_out) ms_out[v] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    PB_HIP(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

template <typename T>
void pos_spmv(hipStream_t s, const ColorOrder& co, int64_t F, int L, const DiluFactors<T>& d, const T* x, const T* b,
              T* y, double* partial) {
    const dim3 g((unsigned)((F + kBlock - 1) / kBlock), (unsigned)L);
    hipLaunchKernelGGL((k_pos_spmv<T, true>), g, dim3(kBlock), 0, s, F, L, co.nbp.p, d.diag.p, d.lat.p, d.vert.p, x,
                       b, y, partial);
}

template <typename Ti, typename To>
void to_pos(hipStream_t s, const ColorOrder& co, int64_t F, int L, const Ti* in, To* out) {
    const dim3 g((unsigned)((F + kBlock - 1) / kBlock), (unsigned)L);
    hipLaunchKernelGGL((k_to_pos<Ti, To>), g, dim3(kBlock), 0, s, F, co.faces.p, in, out);
}
template <typename Ti, typename To>
void from_pos(hipStream_t s, const ColorOrder& co, int64_t F, int L, const Ti* in, To* out) {
    const dim3 g((unsigned)((F + kBlock - 1) / kBlock), (unsigned)L);
    hipLaunchKernelGGL((k_from_pos<Ti, To>), g, dim3(kBlock), 0, s, F, co.faces.p, in, out);
}

template void dilu_factor<double>(hipStream_t, const ColorOrder&, int64_t, int, const int32_t*, const double*,
                                  const uint16_t*, DiluFactors<double>&);
template void dilu_factor<float>(hipStream_t, const ColorOrder&, int64_t, int, const int32_t*, const float*,
                                 const uint16_t*, DiluFactors<float>&);
template void dilu_apply<double>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<double>&, const double*,
                                 double*, bool, bool);
template void dilu_apply<float>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<float>&, const float*,
                                float*, bool, bool);
template void split_apply_batch<double>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<double>&,
                                        const SweepTables&, const double*, double*, const double*, double*,
                                        const double*, int);
template void split_apply_batch<float>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<float>&,
                                       const SweepTables&, const float*, float*, const float*, double*, const double*,
                                       int);
template void split_op<double>(int, hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<double>&, const double*,
                               double*, int);
template void split_op<float>(int, hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<float>&, const float*,
                              float*, int);
template int64_t split_apply_dots<double>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<double>&,
                                          const double*, double*, const double*, double*, int, const double*);
template int64_t split_apply_dots<float>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<float>&,
                                         const float*, float*, const float*, double*, int, const double*);
template int tune_split_mode<double>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<double>&,
                                     const double*, double*, float*);
template int tune_split_mode<float>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<float>&, const float*,
                                    float*, float*);
template bool tune_dilu_xcd<double>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<double>&, const double*,
                                    double*, float*);
template bool tune_dilu_xcd<float>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<float>&, const float*,
                                   float*, float*);
template void pos_spmv<double>(hipStream_t, const ColorOrder&, int64_t, int, const DiluFactors<double>&,
                               const double*, const double*, double*, double*);
template void pos_spmv_variant<double>(hipStream_t, int, const ColorOrder&, int64_t, int, const DiluFactors<double>&,
                                       const double*, double*);
template void pos_spmv_variant<float>(hipStream_t, int, const ColorOrder&, int64_t, int, const DiluFactors<float>&,
                                      const float*, float*);
template int64_t pos_spmv_dots<double>(hipStream_t, int, const ColorOrder&, int64_t, int, const DiluFactors<double>&,
                                       const double*, double*, const double*, double*);
template int64_t pos_spmv_dots<float>(hipStream_t, int, const ColorOrder&, int64_t, int, const DiluFactors<float>&,
                                      const float*, float*, const float*, double*);
template int tune_pos_spmv<double>(hipStream_t, const ColorOrder&, int64_t, int, const DiluFactors<double>&,
                                   const double*, double*, float*);
template int tune_pos_spmv<float>(hipStream_t, const ColorOrder&, int64_t, int, const DiluFactors<float>&,
                                  const float*, float*, float*);
template void pos_spmv<float>(hipStream_t, const ColorOrder&, int64_t, int, const DiluFactors<float>&, const float*,
                              const float*, float*, double*);
template void to_pos<double, double>(hipStream_t, const ColorOrder&, int64_t, int, const double*, double*);
template void to_pos<double, float>(hipStream_t, const ColorOrder&, int64_t, int, const double*, float*);
template void from_pos<double, double>(hipStream_t, const ColorOrder&, int64_t, int, const double*, double*);

}  // namespace pb
