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

// In-register pair update on register bit J.  PRED: some controls sit on
// register bits; the test on the constant p is wavefront-uniform, and the
// update is committed with a select so that no control flow splits the
// register arrays.
template <int R, int AK, int J, bool PRED>
__device__ __forceinline__ void pair_j(double *vr, double *vi, const double *k, uint32_t mro,
                                       uint32_t mrz)
{
#pragma unroll
    for (int p = 0; p < R; p++) {
        if (p & (1 << J)) continue;
        const int q = p | (1 << J);
        if constexpr (!PRED) {
            pair_update(AK, k, vr[p], vi[p], vr[q], vi[q]);
        } else {
            const bool ok = ((uint32_t)p & mro) == mro && !((uint32_t)p & mrz);
            double a = vr[p], b = vi[p], c = vr[q], d = vi[q];
            pair_update(AK, k, a, b, c, d);
            vr[p] = ok ? a : vr[p];
            vi[p] = ok ? b : vi[p];
            vr[q] = ok ? c : vr[q];
            vi[q] = ok ? d : vi[q];
        }
    }
}

template <int R, int RB, int AK>
__device__ __forceinline__ void pair_slot(int slot, double *vr, double *vi, const double *k,
                                          uint32_t mro, uint32_t mrz)
{
    if ((mro | mrz) == 0) {
        switch (slot) {
        case 0: pair_j<R, AK, 0, false>(vr, vi, k, 0, 0); break;
        case 1: if constexpr (RB > 1) pair_j<R, AK, 1, false>(vr, vi, k, 0, 0); break;
        case 2: if constexpr (RB > 2) pair_j<R, AK, 2, false>(vr, vi, k, 0, 0); break;
        case 3: if constexpr (RB > 3) pair_j<R, AK, 3, false>(vr, vi, k, 0, 0); break;
        default: break;
        }
    } else {
        switch (slot) {
        case 0: pair_j<R, AK, 0, true>(vr, vi, k, mro, mrz); break;
        case 1: if constexpr (RB > 1) pair_j<R, AK, 1, true>(vr, vi, k, mro, mrz); break;
        case 2: if constexpr (RB > 2) pair_j<R, AK, 2, true>(vr, vi, k, mro, mrz); break;
        case 3: if constexpr (RB > 3) pair_j<R, AK, 3, true>(vr, vi, k, mro, mrz); break;
        default: break;
        }
    }
}

// Diagonal update of one amplitude; `bit` is the target bit, the two
// multipliers are values (never selected by address).
template <int AK>
__device__ __forceinline__ void diag_one(const double *k, bool bit, double &x, double &y)
{
    if constexpr (AK == AK_NEG) {  // z_proc_chunk: negate the |1> amplitude
        const double nx = -x, ny = -y;
        x = bit ? nx : x;
        y = bit ? ny : y;
    } else if constexpr (AK == AK_PHASE) {  // p_proc_chunk on the |1> amplitude
        const double zr = x, zi = y;
        const double nx = __builtin_fma(zr, k[0], -zi * k[1]);
        const double ny = __builtin_fma(zi, k[0], zr * k[1]);
        x = bit ? nx : x;
        y = bit ? ny : y;
    } else {  // AK_RZ: multiply by diag[bit]
        const double mr = bit ? k[2] : k[0];
        const double mi = bit ? k[3] : k[1];
        const double cr = x, ci = y;
        x = __builtin_fma(cr, mr, -ci * mi);
        y = __builtin_fma(cr, mi, ci * mr);
    }
}

// DSEL: 0 = target on register bit J (bit known per register), 1 = target on
// a lane bit (lanebit), 2 = target outside the tile (ebit, uniform).
template <int R, int AK, int DSEL, int J, bool PRED>
__device__ __forceinline__ void diag_j(double *vr, double *vi, const double *k, bool lanebit,
                                       uint32_t mro, uint32_t mrz)
{
#pragma unroll
    for (int p = 0; p < R; p++) {
        const bool bit = DSEL == 0 ? (((p >> J) & 1) != 0) : lanebit;
        if constexpr (DSEL == 0 && !PRED && (AK == AK_NEG || AK == AK_PHASE)) {
            if (!bit) continue;  // |0> amplitudes are untouched: skipped at compile time
        }
        if constexpr (!PRED) {
            diag_one<AK>(k, bit, vr[p], vi[p]);
        } else {
            const bool ok = ((uint32_t)p & mro) == mro && !((uint32_t)p & mrz);
            double x = vr[p], y = vi[p];
            diag_one<AK>(k, bit, x, y);
            vr[p] = ok ? x : vr[p];
            vi[p] = ok ? y : vi[p];
        }
    }
}

template <int R, int RB, int AK, bool PRED>
__device__ __forceinline__ void diag_sel(double *vr, double *vi, const double *k, int dsel,
                                         int dpos, bool lanebit, uint32_t mro, uint32_t mrz)
{
    if (dsel == 0) {
        switch (dpos) {
        case 0: diag_j<R, AK, 0, 0, PRED>(vr, vi, k, false, mro, mrz); break;
        case 1: if constexpr (RB > 1) diag_j<R, AK, 0, 1, PRED>(vr, vi, k, false, mro, mrz); break;
        case 2: if constexpr (RB > 2) diag_j<R, AK, 0, 2, PRED>(vr, vi, k, false, mro, mrz); break;
        case 3: if constexpr (RB > 3) diag_j<R, AK, 0, 3, PRED>(vr, vi, k, false, mro, mrz); break;
        default: break;
        }
    } else {
        diag_j<R, AK, 1, 0, PRED>(vr, vi, k, lanebit, mro, mrz);
    }
}

template <int R, int RB, int AK>
__device__ __forceinline__ void diag_all(double *vr, double *vi, const double *k, int dsel,
                                         int dpos, bool lanebit, uint32_t mro, uint32_t mrz)
{
    if ((mro | mrz) == 0) diag_sel<R, RB, AK, false>(vr, vi, k, dsel, dpos, lanebit, 0, 0);
    else diag_sel<R, RB, AK, true>(vr, vi, k, dsel, dpos, lanebit, mro, mrz);
}

// Register part of the local index of register p under a layout.
template <int RB>
__device__ __forceinline__ uint32_t reg_part(int p, const int8_t *regbits)
{
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < RB; j++)
        if ((p >> j) & 1) r |= 1u << regbits[j];
    return r;
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
#pragma unroll
    for (int p = 0; p < R; p++) {
        const uint64_t g = base | gt | d.gr[p];
        vr[p] = (double)re[g];
        vi[p] = (double)im[g];
    }
    // lane part of the local index: in the load layout the thread index holds
    // local bits 0..LOGW-1 and register p holds bits LOGW..K-1
    uint32_t lt = tid;

    const int nops = d.nops;
    TileOp cur = ops[0];
    for (int o = 0; o < nops; o++) {
        // issue the next op's scalar loads before working on this one
        const TileOp nxt = ops[o + 1 < nops ? o + 1 : o];
        if (cur.type == TOP_LAYOUT) {
            uint32_t nlt = 0;
#pragma unroll
            for (int i = 0; i < LOGW; i++)
                if ((tid >> i) & 1u) nlt |= 1u << cur.lay.newthr[i];
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) lds[swz(lt | reg_part<RB>(p, cur.lay.oldreg))] = vr[p];
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) vr[p] = lds[swz(nlt | reg_part<RB>(p, cur.lay.newreg))];
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) lds[swz(lt | reg_part<RB>(p, cur.lay.oldreg))] = vi[p];
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) vi[p] = lds[swz(nlt | reg_part<RB>(p, cur.lay.newreg))];
            lt = nlt;
        } else if ((base & cur.eo) == cur.eo && !(base & cur.ez)) {
            // (a control outside the tile that is not met skips the op)
            const uint32_t mlo = cur.mlo, mlz = cur.mlz;
            const bool lane_ok = ((mlo | mlz) == 0) || ((lt & mlo) == mlo && !(lt & mlz));
            if (lane_ok) {
                const uint32_t mro = cur.mro, mrz = cur.mrz;
                const int ak = cur.ak;
                const double *k = cur.c;
                if (cur.type == TOP_PAIR) {
                    const int slot = cur.slot;
                    switch (ak) {
                    case AK_H: pair_slot<R, RB, AK_H>(slot, vr, vi, k, mro, mrz); break;
                    case AK_X: pair_slot<R, RB, AK_X>(slot, vr, vi, k, mro, mrz); break;
                    case AK_Y: pair_slot<R, RB, AK_Y>(slot, vr, vi, k, mro, mrz); break;
                    case AK_RX: pair_slot<R, RB, AK_RX>(slot, vr, vi, k, mro, mrz); break;
                    case AK_RY: pair_slot<R, RB, AK_RY>(slot, vr, vi, k, mro, mrz); break;
                    case AK_U: pair_slot<R, RB, AK_U>(slot, vr, vi, k, mro, mrz); break;
                    case AK_MAT: pair_slot<R, RB, AK_MAT>(slot, vr, vi, k, mro, mrz); break;
                    default: break;
                    }
                } else {
                    int dsel = cur.dsel;
                    const int dpos = cur.dpos;
                    bool lanebit;
                    if (dsel == 2) {  // target outside the tile: the same bit for all
                        lanebit = (base & cur.ebit) != 0;
                        dsel = 1;
                    } else {
                        lanebit = ((lt >> dpos) & 1u) != 0;
                    }
                    switch (ak) {
                    case AK_NEG: diag_all<R, RB, AK_NEG>(vr, vi, k, dsel, dpos, lanebit, mro, mrz); break;
                    case AK_PHASE: diag_all<R, RB, AK_PHASE>(vr, vi, k, dsel, dpos, lanebit, mro, mrz); break;
                    case AK_RZ: diag_all<R, RB, AK_RZ>(vr, vi, k, dsel, dpos, lanebit, mro, mrz); break;
                    default: break;
                    }
                }
            }
        }
        cur = nxt;
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
