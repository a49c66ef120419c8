// spz_tile.hip -- the fused tile-sweep kernel for gfx950.
//
// One workgroup per tile (see spz_tile.h).  The sweep's program is an array
// of TileOps in device memory; every thread walks the same program (the op
// index is wavefront-uniform, so op fields come in through scalar loads).
// Register arrays are only ever indexed by compile-time constants: the
// target's register bit selects a template instance, so the amplitudes stay
// in VGPRs.  Per-gate arithmetic is the same pair_update/diag_update as the
// plain kernels, so a sweep that keeps program order is bit-identical to
// one launch per gate.
#include "spz_internal.h"
#include "spz_tile.h"

namespace spz {

namespace {

// LDS bank swizzle for 8-byte elements: XOR the low 4 bits of the element
// index with two higher nibbles so that power-of-two strides across lanes
// spread over the banks (a bijection inside every aligned block of 16).
__device__ __forceinline__ uint32_t swz(uint32_t l) { return l ^ ((l >> 4) & 15u) ^ ((l >> 8) & 15u); }

template <int R, int AK, int J>
__device__ __forceinline__ void pair_j(double *vr, double *vi, const uint32_t *li,
                                       const double *c, uint32_t lc, uint32_t lz)
{
    if (lc | lz) {
#pragma unroll
        for (int p = 0; p < R; p++) {
            if (p & (1 << J)) continue;
            const int q = p | (1 << J);
            const uint32_t l = li[p];
            if ((l & lc) == lc && !(l & lz)) pair_update(AK, c, vr[p], vi[p], vr[q], vi[q]);
        }
    } else {
#pragma unroll
        for (int p = 0; p < R; p++) {
            if (p & (1 << J)) continue;
            const int q = p | (1 << J);
            pair_update(AK, c, vr[p], vi[p], vr[q], vi[q]);
        }
    }
}

template <int R, int RB, int AK>
__device__ __forceinline__ void pair_slot(int slot, double *vr, double *vi, const uint32_t *li,
                                          const double *c, uint32_t lc, uint32_t lz)
{
    switch (slot) {
    case 0: pair_j<R, AK, 0>(vr, vi, li, c, lc, lz); break;
    case 1: if constexpr (RB > 1) pair_j<R, AK, 1>(vr, vi, li, c, lc, lz); break;
    case 2: if constexpr (RB > 2) pair_j<R, AK, 2>(vr, vi, li, c, lc, lz); break;
    case 3: if constexpr (RB > 3) pair_j<R, AK, 3>(vr, vi, li, c, lc, lz); break;
    case 4: if constexpr (RB > 4) pair_j<R, AK, 4>(vr, vi, li, c, lc, lz); break;
    default: break;
    }
}

template <int R, int AK>
__device__ __forceinline__ void diag_all(double *vr, double *vi, const uint32_t *li,
                                         const double *c, int dbit, int ebitset, uint32_t lc,
                                         uint32_t lz)
{
#pragma unroll
    for (int p = 0; p < R; p++) {
        const uint32_t l = li[p];
        if ((l & lc) != lc || (l & lz)) continue;
        const int bit = dbit >= 0 ? (int)((l >> dbit) & 1u) : ebitset;
        diag_update(AK, c, bit, vr[p], vi[p]);
    }
}

template <typename T, int K, int LOGW>
__global__ void __launch_bounds__(1 << LOGW)
k_tile(T *__restrict__ re, T *__restrict__ im, const TileDesc d, const TileOp *__restrict__ ops)
{
    constexpr int RB = K - LOGW;
    constexpr int R = 1 << RB;
    __shared__ double lds[1 << K];
    const uint32_t tid = threadIdx.x;

    // global bits of the qubits outside S for this tile
    const uint64_t base = insert_zeros((uint64_t)blockIdx.x, d.spos, K);
    uint64_t gt = 0;
#pragma unroll
    for (int i = 0; i < LOGW; i++)
        if ((tid >> i) & 1u) gt |= 1ull << d.spos[i];

    double vr[R], vi[R];
    uint32_t li[R];
#pragma unroll
    for (int p = 0; p < R; p++) {
        const uint64_t g = base | gt | d.gr[p];
        vr[p] = (double)re[g];
        vi[p] = (double)im[g];
        li[p] = ((uint32_t)p << LOGW) | tid;
    }

    const int nops = d.nops;
    for (int o = 0; o < nops; o++) {
        const TileOp *op = ops + o;
        const int type = op->type;
        if (type == TOP_LAYOUT) {
            uint32_t lt = 0;
#pragma unroll
            for (int i = 0; i < LOGW; i++)
                if ((tid >> i) & 1u) lt |= 1u << op->thrbits[i];
            uint32_t nl[R];
#pragma unroll
            for (int p = 0; p < R; p++) {
                uint32_t lr = 0;
#pragma unroll
                for (int j = 0; j < RB; j++)
                    if ((p >> j) & 1) lr |= 1u << op->regbits[j];
                nl[p] = lt | lr;
            }
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) lds[swz(li[p])] = vr[p];
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) vr[p] = lds[swz(nl[p])];
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) lds[swz(li[p])] = vi[p];
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) {
                vi[p] = lds[swz(nl[p])];
                li[p] = nl[p];
            }
            continue;
        }
        const uint64_t eo = op->eo, ez = op->ez;
        if ((base & eo) != eo || (base & ez)) continue;  // control outside the tile not met
        const uint32_t lc = op->lc, lz = op->lz;
        const int ak = op->ak;
        const double *c = op->c;
        if (type == TOP_PAIR) {
            const int slot = op->slot;
            switch (ak) {
            case AK_H: pair_slot<R, RB, AK_H>(slot, vr, vi, li, c, lc, lz); break;
            case AK_X: pair_slot<R, RB, AK_X>(slot, vr, vi, li, c, lc, lz); break;
            case AK_Y: pair_slot<R, RB, AK_Y>(slot, vr, vi, li, c, lc, lz); break;
            case AK_RX: pair_slot<R, RB, AK_RX>(slot, vr, vi, li, c, lc, lz); break;
            case AK_RY: pair_slot<R, RB, AK_RY>(slot, vr, vi, li, c, lc, lz); break;
            case AK_U: pair_slot<R, RB, AK_U>(slot, vr, vi, li, c, lc, lz); break;
            case AK_MAT: pair_slot<R, RB, AK_MAT>(slot, vr, vi, li, c, lc, lz); break;
            default: break;
            }
        } else {
            const int dbit = op->dbit;
            const int ebitset = (base & op->ebit) ? 1 : 0;
            switch (ak) {
            case AK_NEG: diag_all<R, AK_NEG>(vr, vi, li, c, dbit, ebitset, lc, lz); break;
            case AK_PHASE: diag_all<R, AK_PHASE>(vr, vi, li, c, dbit, ebitset, lc, lz); break;
            case AK_RZ: diag_all<R, AK_RZ>(vr, vi, li, c, dbit, ebitset, lc, lz); break;
            default: break;
            }
        }
    }

#pragma unroll
    for (int p = 0; p < R; p++) {
        const uint64_t g = base | gt | d.gr[p];
        re[g] = (T)vr[p];
        im[g] = (T)vi[p];
    }
}

template <typename T, int K>
int launch_k(spz_state_s *s, const TileDesc &d, const TileOp *d_ops)
{
    constexpr int LOGW = K - 4;  // 16 amplitudes per thread
    const uint64_t tiles = s->N >> K;
    if (tiles == 0 || tiles > (1ull << 31) - 1) return fail(SPZ_ERR_ARG, "tile grid out of range");
    const double bytes = 2.0 * (double)s->N * (s->prec == SPZ_F64 ? 16.0 : 8.0);
    LaunchTimer tm(SPZ_K_TILE, bytes);
    hipLaunchKernelGGL((k_tile<T, K, LOGW>), dim3((unsigned)tiles), dim3(1 << LOGW), 0,
                       rt().stream, (T *)s->re, (T *)s->im, d, d_ops);
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

}  // namespace

int launch_tile(spz_state_s *s, int K, const TileDesc &d, const TileOp *d_ops)
{
    if (s->prec == SPZ_F64) {
        switch (K) {
        case 10: return launch_k<double, 10>(s, d, d_ops);
        case 11: return launch_k<double, 11>(s, d, d_ops);
        case 12: return launch_k<double, 12>(s, d, d_ops);
        case 13: return launch_k<double, 13>(s, d, d_ops);
        default: break;
        }
    } else {
        switch (K) {
        case 10: return launch_k<float, 10>(s, d, d_ops);
        case 11: return launch_k<float, 11>(s, d, d_ops);
        case 12: return launch_k<float, 12>(s, d, d_ops);
        case 13: return launch_k<float, 13>(s, d, d_ops);
        default: break;
        }
    }
    return fail(SPZ_ERR_ARG, "unsupported tile width");
}

}  // namespace spz
