// spz_tile_kernel.h -- the fused tile-sweep kernel for gfx950, included by
// spz_tile.hip (the exact arithmetic) and spz_tile_fast.hip (the AUTO
// arithmetic), so that the two instantiations compile in parallel.
#pragma once
//
// One workgroup per tile (see spz_tile.h).  The sweep's program is an array
// of TileOps in device memory; every thread walks the same program (the op
// index is wavefront-uniform, so op fields come in through scalar loads).
// Register arrays are only ever indexed by compile-time constants: the
// target's register bit selects a template instance, so the amplitudes stay
// in VGPRs.
//
// Two rules keep the amplitudes in one register set for the whole program:
//  * dispatch is a chain of independent ifs on one-hot bits the host sets
//    (TileOp::grp, TileOp::sel).  A multi-way branch (a switch, or ifs the
//    optimiser can prove exclusive) merges ~90 paths into one join, and the
//    register allocator then keeps the merged amplitudes apart from the
//    loop-carried ones and copies all of them on every op;
//  * every update is in place: the 2x2 arithmetic is written as inline
//    assembly with tied operands, ordered so that each result overwrites an
//    input whose last use has passed (X and Y are register swaps).
// Without FAST the instructions and their order are those of
// pair_update/diag_update (the reference's per-gate arithmetic, FMA exactly
// where it writes mul_add), so a sweep that keeps program order is
// bit-identical to one launch per gate.  FAST (AUTO sweeps) trades that for
// fewer instructions: H without its scale, rotations and phases as shears,
// Paulis and global phases applied once at the store (the host plans them).
// Layout changes move data only (LDS
// transposes, or in-wavefront lane swaps) and never change a value.
#include <type_traits>

#include "spz_inplace.h"
#include "spz_internal.h"
#include "spz_tile.h"

namespace spz {

namespace {

// The LDS swizzle (tile_swz) is linear, so the per-lane term of an LDS
// address is formed once per layout change and the per-register term is
// wavefront-uniform (scalar), leaving one vector XOR per LDS address.
__device__ __forceinline__ uint32_t swz(uint32_t l) { return tile_swz(l); }

// Dispatch tests are marked unlikely, so each op's body sits out of line and
// the tests that miss (nearly all of them, for every op) fall through untaken.
__device__ __forceinline__ bool hit(uint32_t bits) { return __builtin_expect(bits != 0u, 0); }

// Pseudo-kind of the kernel's template dispatch: an H scaled by k[0] (the
// reference's H when k[0] is 1/sqrt(2)); used by fast sweeps, whose plain
// H ops leave the factor to a later gate.
constexpr int AK_HSCALED = 100;

// FAST selects the arithmetic of AUTO sweeps: RX and RY as two rotations of
// real pairs, three FMAs each (shears; the host gives (-tan(a/2), sin a)), H
// as a sum and a difference whose 1/sqrt(2) the host carries to the store,
// an uncontrolled general 2x2 as a phase, an RY and a phase (AK_USH).
// Without FAST every body is the reference's operation sequence.
template <int AK, bool FAST>
__device__ __forceinline__ void ip_pair(const double *k, double &a, double &b, double &c, double &d)
{
    if constexpr (AK == AK_H) {
        if constexpr (FAST) ip_h_unscaled(a, b, c, d);
        else ip_h(k[0], a, b, c, d);
    } else if constexpr (AK == AK_HSCALED) ip_h(k[0], a, b, c, d);
    else if constexpr (AK == AK_X) { ip_swap(a, c); ip_swap(b, d); }
    else if constexpr (AK == AK_Y) ip_y(a, b, c, d);
    else if constexpr (AK == AK_RX) {
        if constexpr (FAST) ip_rx_shear(k[0], k[1], a, b, c, d);
        else ip_rx(k[0], k[1], a, b, c, d);
    } else if constexpr (AK == AK_RY) {
        if constexpr (FAST) ip_ry_shear(k[0], k[1], a, b, c, d);
        else ip_ry(k[0], k[1], a, b, c, d);
    } else if constexpr (AK == AK_MAT) ip_mat(k, a, b, c, d);
    else if constexpr (AK == AK_USH) {  // psi1 e^(id), RY, psi1 e^(ib): 12 FMAs
        ip_rot<false>(k[0], k[1], c, d);
        ip_ry_shear(k[2], k[3], a, b, c, d);
        ip_rot<false>(k[4], k[5], c, d);
    } else ip_u(k, a, b, c, d);
}

// Lane-bit exchange of one register pair: the halves of x (register bit 0)
// and y (register bit 1) trade places so that register bit and lane bit 4
// (W = 16) or 5 (W = 32) swap roles.  v_permlane*_swap writes both of its
// operands in place.
template <int W>
__device__ __forceinline__ void lane_swap(double &x, double &y)
{
    const uint64_t ux = __builtin_bit_cast(uint64_t, x), uy = __builtin_bit_cast(uint64_t, y);
    uint32_t xl = (uint32_t)ux, xh = (uint32_t)(ux >> 32);
    uint32_t yl = (uint32_t)uy, yh = (uint32_t)(uy >> 32);
    if constexpr (W == 32) {
        const auto l = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
        xl = l[0]; yl = l[1]; xh = h[0]; yh = h[1];
    } else {
        const auto l = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
        xl = l[0]; yl = l[1]; xh = h[0]; yh = h[1];
    }
    x = __builtin_bit_cast(double, (uint64_t)xl | ((uint64_t)xh << 32));
    y = __builtin_bit_cast(double, (uint64_t)yl | ((uint64_t)yh << 32));
}

template <int R, int J, int W>
__device__ __forceinline__ void swap_j(double *vr, double *vi)
{
#pragma unroll
    for (int p = 0; p < R; p++) {
        if (p & (1 << J)) continue;
        lane_swap<W>(vr[p], vr[p | (1 << J)]);
        lane_swap<W>(vi[p], vi[p | (1 << J)]);
    }
}

template <int R, int RB, int W>
__device__ __forceinline__ void swap_sel(uint32_t sel, double *vr, double *vi)
{
    if (hit(sel & 1u)) swap_j<R, 0, W>(vr, vi);
    if constexpr (RB > 1) if (hit(sel & 2u)) swap_j<R, 1, W>(vr, vi);
    if constexpr (RB > 2) if (hit(sel & 4u)) swap_j<R, 2, W>(vr, vi);
    if constexpr (RB > 3) if (hit(sel & 8u)) swap_j<R, 3, W>(vr, vi);
}

// ---------------------------------------------------------------------------
// Dispatch (one-hot bits set by the host; see TileOp::grp)
// ---------------------------------------------------------------------------

// Pair update on register bit J.  PRED: controls on register bits, as the
// host-computed mask pm of the registers that meet them (uniform); controls
// on lane bits are one divergent branch around the whole op (see k_tile).
template <int R, int AK, int J, bool PRED, bool FAST>
__device__ __forceinline__ void pair_j(double *vr, double *vi, const double *k, uint32_t pm)
{
#pragma unroll
    for (int p = 0; p < R; p++) {
        if (p & (1 << J)) continue;
        const int q = p | (1 << J);
        if constexpr (!PRED) {
            ip_pair<AK, FAST>(k, vr[p], vi[p], vr[q], vi[q]);
        } else {
            if ((pm >> p) & 1u)
                ip_pair<AK, FAST>(k, vr[p], vi[p], vr[q], vi[q]);
        }
    }
}

template <int R, int RB, int AK, bool PRED, bool FAST>
__device__ __forceinline__ void pair_sel(uint32_t sel, double *vr, double *vi, const double *k,
                                         uint32_t pm)
{
    if (hit(sel & 1u)) pair_j<R, AK, 0, PRED, FAST>(vr, vi, k, pm);
    if constexpr (RB > 1) if (hit(sel & 2u)) pair_j<R, AK, 1, PRED, FAST>(vr, vi, k, pm);
    if constexpr (RB > 2) if (hit(sel & 4u)) pair_j<R, AK, 2, PRED, FAST>(vr, vi, k, pm);
    if constexpr (RB > 3) if (hit(sel & 8u)) pair_j<R, AK, 3, PRED, FAST>(vr, vi, k, pm);
}

template <int R, int RB, bool PRED, bool FAST>
__device__ __forceinline__ void pair_kinds(uint32_t g, uint32_t sel, double *vr, double *vi,
                                           const double *k, uint32_t pm)
{
    if (hit(g & (TG_H | TG_X | TG_Y | TG_HS))) {
        if (hit(g & TG_H)) pair_sel<R, RB, AK_H, PRED, FAST>(sel, vr, vi, k, pm);
        if (hit(g & TG_X)) pair_sel<R, RB, AK_X, PRED, FAST>(sel, vr, vi, k, pm);
        if (hit(g & TG_Y)) pair_sel<R, RB, AK_Y, PRED, FAST>(sel, vr, vi, k, pm);
        if constexpr (FAST)
            if (hit(g & TG_HS)) pair_sel<R, RB, AK_HSCALED, PRED, FAST>(sel, vr, vi, k, pm);
    }
    if (hit(g & (TG_RX | TG_RY | TG_U | TG_MAT | TG_USH))) {
        if (hit(g & TG_RX)) pair_sel<R, RB, AK_RX, PRED, FAST>(sel, vr, vi, k, pm);
        if (hit(g & TG_RY)) pair_sel<R, RB, AK_RY, PRED, FAST>(sel, vr, vi, k, pm);
        if (hit(g & TG_U)) pair_sel<R, RB, AK_U, PRED, FAST>(sel, vr, vi, k, pm);
        if constexpr (FAST) {
            if (hit(g & TG_MAT)) pair_sel<R, RB, AK_MAT, PRED, FAST>(sel, vr, vi, k, pm);
            if constexpr (!PRED)  // (uncontrolled only)
                if (hit(g & TG_USH)) pair_sel<R, RB, AK_USH, PRED, FAST>(sel, vr, vi, k, pm);
        }
    }
}

// Diagonal update of one amplitude whose target bit is known at compile
// time (BITK) or given per lane / per tile (bit, with per-lane multipliers
// mr, mi for RZ).
// FAST: the phases are rotations of (re, im) as three shears, the host
// giving (-tan(b/2), sin b) for a phase e^(ib), |b| <= pi/2.
template <int AK, bool FAST>
__device__ __forceinline__ void diag_known(const double *k, bool bit, double &x, double &y)
{
    if constexpr (AK == AK_NEG) {
        if (bit) { x = -x; y = -y; }
    } else if constexpr (AK == AK_PHASE) {
        if constexpr (FAST) {
            if (bit) ip_rot<false>(k[0], k[1], x, y);
        } else {
            if (bit) ip_phase(k[0], k[1], x, y);
        }
    } else {  // AK_RZ: diag[bit]
        if constexpr (FAST) {
            if (bit) ip_rot<false>(k[2], k[3], x, y);
            else ip_rot<false>(k[0], k[1], x, y);
        } else {
            if (bit) ip_rz<false>(k[2], k[3], x, y);
            else ip_rz<false>(k[0], k[1], x, y);
        }
    }
}

// Diagonal target on register bit J.
template <int R, int AK, int J, bool PRED, bool FAST>
__device__ __forceinline__ void diag_j(double *vr, double *vi, const double *k, uint32_t pm)
{
#pragma unroll
    for (int p = 0; p < R; p++) {
        const bool bit = ((p >> J) & 1) != 0;
        if constexpr (AK == AK_NEG || AK == AK_PHASE)
            if (!bit) continue;  // |0> amplitudes are untouched
        if constexpr (!PRED) {
            diag_known<AK, FAST>(k, bit, vr[p], vi[p]);
        } else {
            if ((pm >> p) & 1u)
                diag_known<AK, FAST>(k, bit, vr[p], vi[p]);
        }
    }
}

template <int R, int RB, int AK, bool PRED, bool FAST>
__device__ __forceinline__ void diag_sel(uint32_t sel, double *vr, double *vi, const double *k,
                                         uint32_t pm)
{
    if (hit(sel & 1u)) diag_j<R, AK, 0, PRED, FAST>(vr, vi, k, pm);
    if constexpr (RB > 1) if (hit(sel & 2u)) diag_j<R, AK, 1, PRED, FAST>(vr, vi, k, pm);
    if constexpr (RB > 2) if (hit(sel & 4u)) diag_j<R, AK, 2, PRED, FAST>(vr, vi, k, pm);
    if constexpr (RB > 3) if (hit(sel & 8u)) diag_j<R, AK, 3, PRED, FAST>(vr, vi, k, pm);
}

// Diagonal target on a lane bit or outside the tile: `bit` per lane (or per
// tile); the RZ multipliers are selected once per op.
template <int R, int AK, bool PRED, bool FAST>
__device__ __forceinline__ void diag_lane(double *vr, double *vi, const double *k, bool bit,
                                          uint32_t pm)
{
    if constexpr (FAST && AK != AK_NEG) {
        // one rotation per amplitude with per-lane shear coefficients (a
        // phase on |1> turns the |0> lanes by zero: exact)
        const double k1 = AK == AK_RZ ? (bit ? k[2] : k[0]) : (bit ? k[0] : 0.0);
        const double k2 = AK == AK_RZ ? (bit ? k[3] : k[1]) : (bit ? k[1] : 0.0);
#pragma unroll
        for (int p = 0; p < R; p++) {
            if constexpr (!PRED) {
                ip_rot<true>(k1, k2, vr[p], vi[p]);
            } else {
                if ((pm >> p) & 1u)
                    ip_rot<true>(k1, k2, vr[p], vi[p]);
            }
        }
    } else if constexpr (AK == AK_RZ) {
        const double mr = bit ? k[2] : k[0];
        const double mi = bit ? k[3] : k[1];
#pragma unroll
        for (int p = 0; p < R; p++) {
            if constexpr (!PRED) {
                ip_rz<true>(mr, mi, vr[p], vi[p]);
            } else {
                if ((pm >> p) & 1u)
                    ip_rz<true>(mr, mi, vr[p], vi[p]);
            }
        }
    } else if constexpr (AK == AK_NEG) {
        // multiply by -1 or +1 per lane: exact (-x, including the sign of zero)
        const double sg = bit ? -1.0 : 1.0;
#pragma unroll
        for (int p = 0; p < R; p++) {
            if constexpr (!PRED) {
                ip_scale(sg, vr[p], vi[p]);
            } else {
                if ((pm >> p) & 1u)
                    ip_scale(sg, vr[p], vi[p]);
            }
        }
    } else if (bit) {
#pragma unroll
        for (int p = 0; p < R; p++) {
            if constexpr (!PRED) {
                diag_known<AK, false>(k, true, vr[p], vi[p]);
            } else {
                if ((pm >> p) & 1u)
                    diag_known<AK, false>(k, true, vr[p], vi[p]);
            }
        }
    }
}

template <int R, int RB, bool PRED, bool FAST>
__device__ __forceinline__ void diag_reg_kinds(uint32_t g, uint32_t sel, double *vr, double *vi,
                                               const double *k, uint32_t pm)
{
    if (hit(g & TG_NEG)) diag_sel<R, RB, AK_NEG, PRED, FAST>(sel, vr, vi, k, pm);
    if (hit(g & TG_PHASE)) diag_sel<R, RB, AK_PHASE, PRED, FAST>(sel, vr, vi, k, pm);
    if (hit(g & TG_RZ)) diag_sel<R, RB, AK_RZ, PRED, FAST>(sel, vr, vi, k, pm);
}

template <int R, bool PRED, bool FAST>
__device__ __forceinline__ void diag_lane_kinds(uint32_t g, double *vr, double *vi,
                                                const double *k, bool bit, uint32_t pm)
{
    if (hit(g & TG_NEG)) diag_lane<R, AK_NEG, PRED, FAST>(vr, vi, k, bit, pm);
    if (hit(g & TG_PHASE)) diag_lane<R, AK_PHASE, PRED, FAST>(vr, vi, k, bit, pm);
    if (hit(g & TG_RZ)) diag_lane<R, AK_RZ, PRED, FAST>(vr, vi, k, bit, pm);
}

// x with its sign bit flipped where bit 31 of m is set (exact, like -x).
__device__ __forceinline__ double flip_sign(double x, uint32_t m)
{
    return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) ^ ((uint64_t)m << 32));
}

// Controls on lane bits, evaluated inside each predicated path (the opaque
// copy keeps the compiler from hoisting the test into every op).
__device__ __forceinline__ bool lane_pass(uint32_t lt, uint32_t mlo, uint32_t mlz)
{
    asm volatile("" : "+v"(lt));
    return (lt & mlo) == mlo && !(lt & mlz);
}

// Sum of the single-bit 64-bit terms t[i] over the set bits i of x (the
// bits i < NB): the per-lane part of a global index.  Each bit becomes an
// all-ones or all-zero word (one signed bit-field extract) that masks the
// halves of its term into the sum (one and-or each).
template <int NB>
__device__ __forceinline__ uint64_t lane_offset(uint32_t x, const uint64_t *t)
{
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const uint32_t m = (uint32_t)((int32_t)(x << (31 - i)) >> 31);
        lo |= m & (uint32_t)t[i];
        hi |= m & (uint32_t)(t[i] >> 32);
    }
    return ((uint64_t)hi << 32) | lo;
}

// Register budget: asking for 6 waves per SIMD makes the compiler fit the
// kernel in ~108 VGPRs without spills (4 resident waves per SIMD instead of 3
// at its unconstrained 138); the extra wave lets one workgroup's loads and
// stores overlap the others' gate arithmetic.
template <typename T, int K, int LOGW, bool FAST>
__global__ void __launch_bounds__(1 << LOGW) __attribute__((amdgpu_waves_per_eu(6, 6)))
k_tile(T *__restrict__ re, T *__restrict__ im, const TileDesc d, const TileOp *__restrict__ ops)
{
    constexpr int RB = K - LOGW;
    constexpr int R = 1 << RB;
    static_assert(RB <= kTileLayoutRB, "a layout change describes at most 4 register-index bits");
    __shared__ double lds[1 << K];
    const uint32_t tid = threadIdx.x;

    // global bits of the qubits outside S for this tile
    const uint64_t base = insert_zeros((uint64_t)blockIdx.x, d.spos, K);
    // thread part of the global index and of the local index (lane part)
    // (branch-free: the uniform table reads are issued together)
    uint32_t lt = 0;
#pragma unroll
    for (int i = 0; i < LOGW; i++) lt |= ((tid >> i) & 1u) << d.tin[i];
    const uint64_t gt = lane_offset<LOGW>(tid, d.tg);

    // base, the register terms gr[p] and the thread part gt have disjoint
    // bits, so an index is their sum: one per-lane pointer for the tile and
    // thread part, plus the uniform register term (launch_k hands gr[] and
    // gs[] over as byte offsets): one vector add per access
    double vr[R], vi[R];
    const char *lr = reinterpret_cast<const char *>(re + (base + gt));
    const char *li = reinterpret_cast<const char *>(im + (base + gt));
#pragma unroll
    for (int p = 0; p < R; p++) {
        vr[p] = (double)*reinterpret_cast<const T *>(lr + d.gr[p]);
        vi[p] = (double)*reinterpret_cast<const T *>(li + d.gr[p]);
    }

    const uint64_t ope = reinterpret_cast<uint64_t>(ops + d.nops);
    for (uint64_t opa = reinterpret_cast<uint64_t>(ops); opa != ope; opa += sizeof(TileOp)) {
        // by value: the op's fields come into scalar registers, so selects
        // between its coefficients are value selects, never per-lane loads.
        // The opaque address read through the constant address space gives
        // scalar loads at immediate offsets of one base that advances by one
        // add per op (not strength-reduced bases per field).
        asm volatile("" : "+s"(opa));
        TileOp cur;
        {
            const __attribute__((address_space(4))) uint64_t *q =
                reinterpret_cast<const __attribute__((address_space(4))) uint64_t *>(opa);
            uint64_t w[sizeof(TileOp) / 8];
#pragma unroll
            for (int i = 0; i < (int)(sizeof(TileOp) / 8); i++) w[i] = q[i];
            __builtin_memcpy(&cur, w, sizeof(TileOp));
        }
        const uint32_t g = cur.grp;
        if (hit(g & TG_ANYLAYOUT)) {
        if (hit(g & (TG_SW16 | TG_SW32))) {
            // in-wavefront change: one register-index bit trades places with
            // lane bit 4 or 5; the lane part of the local index follows
            if (hit(g & TG_SW32)) swap_sel<R, RB, 32>(cur.sel, vr, vi);
            if (hit(g & TG_SW16)) swap_sel<R, RB, 16>(cur.sel, vr, vi);
            const uint32_t tb = (g & TG_SW32) ? 5u : 4u;
            lt = (lt & ~cur.mlo) | (((tid >> tb) & 1u) ? cur.mlz : 0u);
        }
        if (hit(g & TG_LAYOUT)) {
            uint32_t nlt = 0;
#pragma unroll
            for (int i = 0; i < LOGW; i++) nlt |= ((tid >> i) & 1u) << cur.lay.newthr[i];
            // byte offsets into lds: lane term per lane, register term uniform
            uint32_t so = swz(lt) << 3, sn = swz(nlt) << 3;
            // (keeps the compiler from factoring the shift back out of the XORs)
            asm volatile("" : "+v"(so), "+v"(sn));
            // register terms: linear in p, one XOR from p without its lowest bit
            uint32_t ro[R], rn[R];
            ro[0] = 0;
            rn[0] = 0;
#pragma unroll
            for (int p = 1; p < R; p++) {
                const int j = __builtin_ctz(p);
                ro[p] = ro[p & (p - 1)] ^ cur.lay.rob[j];
                rn[p] = rn[p & (p - 1)] ^ cur.lay.rnb[j];
            }
            char *lb = reinterpret_cast<char *>(lds);
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) *reinterpret_cast<double *>(lb + (so ^ ro[p])) = vr[p];
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) vr[p] = *reinterpret_cast<const double *>(lb + (sn ^ rn[p]));
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) *reinterpret_cast<double *>(lb + (so ^ ro[p])) = vi[p];
            __syncthreads();
#pragma unroll
            for (int p = 0; p < R; p++) vi[p] = *reinterpret_cast<const double *>(lb + (sn ^ rn[p]));
            lt = nlt;
        }
        }
        // (a control outside the tile that is not met skips the op)
        if ((base & cur.em) == cur.eo) {
            const uint32_t sel = cur.sel, pm = cur.pm;
            const double *k = cur.c;
            // (two group tests, so an op skips the other family with one test)
            if (hit(g & (TG_PAIR | TG_PAIRP))) {
                if (hit(g & TG_PAIR)) pair_kinds<R, RB, false, FAST>(g, sel, vr, vi, k, 0);
                // predicated ops: the lane part of the controls is one divergent
                // branch around the op, the register part a uniform test per register
                if (hit(g & TG_PAIRP))
                    if (lane_pass(lt, cur.mlo, cur.mlz))
                        pair_kinds<R, RB, true, FAST>(g, sel, vr, vi, k, pm);
            }
            if (hit(g & (TG_DREG | TG_DREGP | TG_DLANE | TG_DLANEP))) {
                if (hit(g & TG_DREG)) diag_reg_kinds<R, RB, false, FAST>(g, sel, vr, vi, k, 0);
                if (hit(g & TG_DREGP))
                    if (lane_pass(lt, cur.mlo, cur.mlz))
                        diag_reg_kinds<R, RB, true, FAST>(g, sel, vr, vi, k, pm);
                if (hit(g & (TG_DLANE | TG_DLANEP))) {
                    // target on a lane bit (dsel 1) or outside the tile (dsel 2)
                    const bool bit = cur.dsel == 2 ? (base & cur.ebit) != 0
                                                   : ((lt >> cur.dpos) & 1u) != 0;
                    if (hit(g & TG_DLANE)) diag_lane_kinds<R, false, FAST>(g, vr, vi, k, bit, 0);
                    if (hit(g & TG_DLANEP))
                        if (lane_pass(lt, cur.mlo, cur.mlz))
                            diag_lane_kinds<R, true, FAST>(g, vr, vi, k, bit, pm);
                }
            }
        }
    }

    // store layout: the lane part is where the last layout change left it;
    // fast sweeps apply their global factor and Pauli frame here (see
    // TileDesc): the X part moves each amplitude to its flipped index (lane
    // bits here, register bits folded into gs[]), the Z part and the phase
    // set its sign, and an imaginary phase stores (-im, re)
    const uint32_t lphi = lt;
    if constexpr (FAST) {
        lt ^= d.fxl;
        if (d.fflags & TF_PHASE) {  // the sweep's global factor (phases, H scale)
#pragma unroll
            for (int p = 0; p < R; p++) ip_cmul(d.gre, d.gim, vr[p], vi[p]);
        }
    }
    uint64_t gsb[K];
#pragma unroll
    for (int b = 0; b < K; b++) gsb[b] = 1ull << d.spos[b];
    const uint64_t gs = lane_offset<K>(lt, gsb);
    char *sr = reinterpret_cast<char *>(re + (base + gs));
    char *si = reinterpret_cast<char *>(im + (base + gs));
    auto store = [&](auto sign, auto odd) {
        uint32_t sl = 0;  // bit 31: odd parity of the lane and tile parts on z
        if constexpr (decltype(sign)::value) {
            sl = (uint32_t)(__builtin_popcount(lphi & d.fzl) +
                            __builtin_popcountll(base & d.fzg)) << 31;
            // (one word per lane: the per-register part below is a uniform
            // word XORed in, not a shift folded into every register's mask)
            asm volatile("" : "+v"(sl));
        }
#pragma unroll
        for (int p = 0; p < R; p++) {
            double a = vr[p], b = vi[p];
            if constexpr (decltype(odd)::value) {
                a = vi[p];
                b = vr[p];
            }
            if constexpr (decltype(sign)::value) {
                const uint32_t u = (d.fzr << (31 - p)) & 0x80000000u;
                a = flip_sign(a, sl ^ (decltype(odd)::value ? u ^ 0x80000000u : u));
                b = flip_sign(b, sl ^ u);
            }
            *reinterpret_cast<T *>(sr + d.gs[p]) = (T)a;
            *reinterpret_cast<T *>(si + d.gs[p]) = (T)b;
        }
    };
    using F = std::false_type;
    using Tr = std::true_type;
    if constexpr (FAST) {
        if (d.fflags & TF_ODD) store(Tr{}, Tr{});
        else if (d.fflags & TF_SIGN) store(Tr{}, F{});
        else store(F{}, F{});
    } else {
        store(F{}, F{});
    }
}

template <typename T, int K, bool FAST>
int launch_k(spz_state_s *s, const TileDesc &d, const TileOp *d_ops)
{
    constexpr int LOGW = K - 4;  // 16 amplitudes per thread
    const uint64_t tiles = s->N >> K;
    if (tiles == 0 || tiles > (1ull << 31) - 1) return fail(SPZ_ERR_ARG, "tile grid out of range");
    const double bytes = 2.0 * (double)s->N * (s->prec == SPZ_F64 ? 16.0 : 8.0);
    // the kernel takes the register terms of its load and store indexes as
    // byte offsets (one uniform add per access, no scaling)
    TileDesc db = d;
    for (int p = 0; p < (1 << (K - LOGW)); p++) {
        db.gr[p] *= sizeof(T);
        db.gs[p] *= sizeof(T);
    }
    LaunchTimer tm(SPZ_K_TILE, bytes);
    hipLaunchKernelGGL((k_tile<T, K, LOGW, FAST>), dim3((unsigned)tiles), dim3(1 << LOGW), 0,
                       rt().stream, (T *)s->re, (T *)s->im, db, d_ops);
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

template <bool FAST>
int launch_fast(spz_state_s *s, int K, const TileDesc &d, const TileOp *d_ops)
{
    if (s->prec == SPZ_F64) {
        switch (K) {
        case 10: return launch_k<double, 10, FAST>(s, d, d_ops);
        case 11: return launch_k<double, 11, FAST>(s, d, d_ops);
        case 12: return launch_k<double, 12, FAST>(s, d, d_ops);
        case 13: return launch_k<double, 13, FAST>(s, d, d_ops);
        default: break;
        }
    } else {
        switch (K) {
        case 10: return launch_k<float, 10, FAST>(s, d, d_ops);
        case 11: return launch_k<float, 11, FAST>(s, d, d_ops);
        case 12: return launch_k<float, 12, FAST>(s, d, d_ops);
        case 13: return launch_k<float, 13, FAST>(s, d, d_ops);
        default: break;
        }
    }
    return fail(SPZ_ERR_ARG, "unsupported tile width");
}

}  // namespace

}  // namespace spz
