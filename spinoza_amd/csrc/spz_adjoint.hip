// spz_adjoint.hip -- adjoint differentiation of E = <psi|H|psi> (no counterpart
// in the reference; built beside the gate path).
//
//   psi_k = U_k ... U_1 psi_0,  lambda_k = U_{k+1}^dag ... U_L^dag H psi_L.
//   For U_k = exp(-i theta G / 2) (RX, RY, RZ with G = X, Y, Z; controls
//   project onto the control subspace), dE/dtheta_k = Im <lambda_k| G~ |psi_k>,
//   and for P(theta) = diag(1, e^{i theta}), dE/dtheta_k = -2 Im <lambda_k|Pi_1~|psi_k>
//   (G~, Pi_1~: the operator restricted to amplitudes whose controls are met).
//
// Reverse sweep.  Consecutive ops whose targets fit in 3 qubits form a group
// that costs one pass over psi and lambda: per element (the 8 amplitudes that
// differ in those qubits) each op adds its gradient term (TwoSum-compensated
// FP64 per thread, a halving tree per block, one more block over the block
// partials, written to a device gradient array) and then applies U_k^dag to
// both states.  Runs of 4 or more parameter-free ops are undone with the
// fused tile executor, once on psi and once on lambda.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

#include "spz_inplace.h"
#include "spz_internal.h"
#include "spz_tile.h"

namespace spz {

int execute_ops(spz_state_s *s, const spz_op *ops, uint64_t n_ops, spz_tracker *trk,
                const spz_exec_opts *opts, uint8_t *outcomes);

namespace {

constexpr int kT = 256;

enum GradKind { GK_NONE = 0, GK_X = 1, GK_Y = 2, GK_Z = 3, GK_P = 4 };

__device__ __forceinline__ void two_sum_d(double a, double b, double &s, double &e)
{
    double ss = a + b;
    double bb = ss - a;
    double aa = ss - bb;
    e = (a - aa) + (b - bb);
    s = ss;
}

// Im(conj(a) b) for a = (ar, ai), b = (br, bi)
__device__ __forceinline__ double im_cj(double ar, double ai, double br, double bi)
{
    return ar * bi - ai * br;
}
__device__ __forceinline__ double re_cj(double ar, double ai, double br, double bi)
{
    return ar * br + ai * bi;
}

// A reverse-sweep group: up to kGMaxOps consecutive ops (taken in reverse
// order) whose targets lie in a set Q of at most kGMaxQ qubits.  One pass
// visits every element (the 2^|Q| amplitudes that differ only in the Q bits,
// with the common control bits fixed) and applies each op's inverse to both
// psi and lambda, after adding the op's gradient term for that element.
constexpr int kGMaxOps = 8;
constexpr int kGMaxQ = 3;
constexpr int kGPerThread = 4;  // elements per thread

struct GroupOp {
    int32_t ak, gk, tq;       // arithmetic, gradient kind, index of the target in Q
    uint32_t kb;              // one-hot kind bit (TileGroup TG_H ... TG_RZ) of ak
    uint64_t cm, ones;        // control mask and the control bits that must be 1
    double c[8];              // inverse-gate coefficients
};

struct GroupDesc {
    int8_t q[kGMaxQ + 1];  // the qubits of Q
    int32_t nop;
    uint64_t base_ones;    // common controls that must be 1 (fixed in every element)
    GroupOp op[kGMaxOps];
};

__device__ __forceinline__ double grad_term(int gk, double la, double lb, double lc, double ld,
                                            double a, double b, double c, double d)
{
    if (gk == GK_X)  // Im <l|X|p> = Im(l0* p1) + Im(l1* p0)
        return im_cj(la, lb, c, d) + im_cj(lc, ld, a, b);
    if (gk == GK_Y)  // Im <l|Y|p> = Re(l1* p0) - Re(l0* p1)
        return re_cj(lc, ld, a, b) - re_cj(la, lb, c, d);
    if (gk == GK_Z)  // Im <l|Z|p> = Im(l0* p0) - Im(l1* p1)
        return im_cj(la, lb, a, b) - im_cj(lc, ld, c, d);
    return -2.0 * im_cj(lc, ld, c, d);  // P: -2 Im <l|Pi_1|p>
}

// The inverse gate on one pair, in place (the kind is a compile-time constant).
template <int AK>
__device__ __forceinline__ void inv_pair(const double *k, double &a, double &b, double &c, double &d)
{
    if constexpr (AK == AK_H) ip_h(k[0], a, b, c, d);
    else if constexpr (AK == AK_X) { ip_swap(a, c); ip_swap(b, d); }
    else if constexpr (AK == AK_Y) ip_y(a, b, c, d);
    else if constexpr (AK == AK_RX) ip_rx(k[0], k[1], a, b, c, d);
    else if constexpr (AK == AK_RY) ip_ry(k[0], k[1], a, b, c, d);
    else if constexpr (AK == AK_U) ip_u(k, a, b, c, d);
    else if constexpr (AK == AK_NEG) { c = -c; d = -d; }
    else if constexpr (AK == AK_PHASE) ip_phase(k[0], k[1], c, d);
    else { ip_rz<false>(k[0], k[1], a, b); ip_rz<false>(k[2], k[3], c, d); }  // AK_RZ
}

// Apply op (target = Q[TQ]) to the pairs of one element, adding the gradient
// terms of the pairs whose controls are met to (s, e).
template <int M, int TQ, int AK>
__device__ __forceinline__ void group_apply(const GroupOp &op, uint64_t base, const uint64_t *off,
                                            double *ar, double *ai, double *lr, double *li,
                                            double &s, double &e)
{
#pragma unroll
    for (int x = 0; x < (1 << M); x++) {
        if (x & (1 << TQ)) continue;
        const int y = x | (1 << TQ);
        if (((base | off[x]) & op.cm) != op.ones) continue;
        if (op.gk != GK_NONE) {
            const double t = grad_term(op.gk, lr[x], li[x], lr[y], li[y], ar[x], ai[x], ar[y], ai[y]);
            double ns, er;
            two_sum_d(s, t, ns, er);
            s = ns;
            e = e + er;
        }
        inv_pair<AK>(op.c, ar[x], ai[x], ar[y], ai[y]);
        inv_pair<AK>(op.c, lr[x], li[x], lr[y], li[y]);
    }
}

// Dispatch on the op's kind through independent tests of one-hot bits (a
// multi-way branch would merge the register arrays of every path and copy them).
template <int M, int TQ>
__device__ __forceinline__ void group_kinds(const GroupOp &op, uint64_t base, const uint64_t *off,
                                            double *ar, double *ai, double *lr, double *li,
                                            double &s, double &e)
{
    const uint32_t kb = op.kb;
    if (kb & TG_H) group_apply<M, TQ, AK_H>(op, base, off, ar, ai, lr, li, s, e);
    if (kb & TG_X) group_apply<M, TQ, AK_X>(op, base, off, ar, ai, lr, li, s, e);
    if (kb & TG_Y) group_apply<M, TQ, AK_Y>(op, base, off, ar, ai, lr, li, s, e);
    if (kb & TG_RX) group_apply<M, TQ, AK_RX>(op, base, off, ar, ai, lr, li, s, e);
    if (kb & TG_RY) group_apply<M, TQ, AK_RY>(op, base, off, ar, ai, lr, li, s, e);
    if (kb & TG_U) group_apply<M, TQ, AK_U>(op, base, off, ar, ai, lr, li, s, e);
    if (kb & TG_NEG) group_apply<M, TQ, AK_NEG>(op, base, off, ar, ai, lr, li, s, e);
    if (kb & TG_PHASE) group_apply<M, TQ, AK_PHASE>(op, base, off, ar, ai, lr, li, s, e);
    if (kb & TG_RZ) group_apply<M, TQ, AK_RZ>(op, base, off, ar, ai, lr, li, s, e);
}

// partials[(j * gridDim.x + block) * 2 + {0, 1}]: compensated block sum of
// op j's gradient terms (ops without a gradient write nothing).
template <typename T, int M>
__global__ void __launch_bounds__(kT)
k_adj_group(T *__restrict__ pr, T *__restrict__ pi, T *__restrict__ qr, T *__restrict__ qi,
            uint64_t nel, PosList pl, GroupDesc gd, double *__restrict__ partials)
{
    __shared__ double S[kGMaxOps][kT], C[kGMaxOps][kT];
    const int tid = threadIdx.x;
    const int nop = gd.nop;
    for (int j = 0; j < nop; j++) {
        S[j][tid] = 0.0;
        C[j][tid] = 0.0;
    }
    uint64_t off[1 << M];
#pragma unroll
    for (int x = 0; x < (1 << M); x++) {
        uint64_t o = 0;
#pragma unroll
        for (int b = 0; b < M; b++)
            if (x & (1 << b)) o |= 1ull << gd.q[b];
        off[x] = o;
    }
    // blocks stride over chunks of kT * kGPerThread elements
    const uint64_t chunk = (uint64_t)kT * kGPerThread;
    for (uint64_t block0 = (uint64_t)blockIdx.x * chunk; block0 < nel;
         block0 += (uint64_t)gridDim.x * chunk) {
        for (int w = 0; w < kGPerThread; w++) {
            const uint64_t u = block0 + (uint64_t)w * kT + tid;
            if (u >= nel) break;
            const uint64_t base = insert_zeros(u, pl.pos, pl.np) | gd.base_ones;
            double ar[1 << M], ai[1 << M], lr[1 << M], li[1 << M];
#pragma unroll
            for (int x = 0; x < (1 << M); x++) {
                const uint64_t i = base | off[x];
                ar[x] = pr[i]; ai[x] = pi[i]; lr[x] = qr[i]; li[x] = qi[i];
            }
            for (int j = 0; j < nop; j++) {
                const GroupOp &op = gd.op[j];
                double s = S[j][tid], e = C[j][tid];
                const uint32_t tqb = 1u << op.tq;  // one-hot, tested independently
                if (tqb & 1u) group_kinds<M, 0>(op, base, off, ar, ai, lr, li, s, e);
                if constexpr (M > 1) if (tqb & 2u) group_kinds<M, 1>(op, base, off, ar, ai, lr, li, s, e);
                if constexpr (M > 2) if (tqb & 4u) group_kinds<M, 2>(op, base, off, ar, ai, lr, li, s, e);
                S[j][tid] = s;
                C[j][tid] = e;
            }
#pragma unroll
            for (int x = 0; x < (1 << M); x++) {
                const uint64_t i = base | off[x];
                pr[i] = (T)ar[x]; pi[i] = (T)ai[x]; qr[i] = (T)lr[x]; qi[i] = (T)li[x];
            }
        }
    }
    __syncthreads();
    for (int h = kT / 2; h >= 1; h >>= 1) {
        if (tid < h)
            for (int j = 0; j < nop; j++) {
                double ns, e;
                two_sum_d(S[j][tid], S[j][tid + h], ns, e);
                C[j][tid] = (C[j][tid] + C[j][tid + h]) + e;
                S[j][tid] = ns;
            }
        __syncthreads();
    }
    if (tid < nop && gd.op[tid].gk != GK_NONE) {
        const uint64_t k = ((uint64_t)tid * gridDim.x + blockIdx.x) * 2;
        partials[k] = S[tid][0];
        partials[k + 1] = C[tid][0];
    }
}

// Block j reduces op j's nb block partials into grad[idx.i[j]].
struct GradIdx {
    int64_t i[kGMaxOps];
};

__global__ void __launch_bounds__(kT)
k_adj_finish(const double *__restrict__ partials, uint64_t nb, GradIdx idx,
             double *__restrict__ grad)
{
    __shared__ double S[kT], C[kT];
    const int tid = threadIdx.x;
    const double *p = partials + (uint64_t)blockIdx.x * nb * 2;
    double s = 0.0, c = 0.0;
    for (uint64_t j = tid; j < nb; j += kT) {
        double ns, e;
        two_sum_d(s, p[2 * j], ns, e);
        c = (c + p[2 * j + 1]) + e;
        s = ns;
    }
    S[tid] = s;
    C[tid] = c;
    __syncthreads();
    for (int h = kT / 2; h >= 1; h >>= 1) {
        if (tid < h) {
            double ns, e;
            two_sum_d(S[tid], S[tid + h], ns, e);
            C[tid] = (C[tid] + C[tid + h]) + e;
            S[tid] = ns;
        }
        __syncthreads();
    }
    if (tid == 0 && idx.i[blockIdx.x] >= 0) grad[idx.i[blockIdx.x]] = S[0] + C[0];
}

// lambda = H psi, H = sum_{i<n-1} Z_i Z_{i+1} + hx sum_i X_i (same operation
// order as the reduction term used for the energy).
template <typename T>
__global__ void __launch_bounds__(kT)
k_ising_apply(const T *__restrict__ re, const T *__restrict__ im, T *__restrict__ lr,
              T *__restrict__ li, uint64_t N, int n, double hx)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += stride) {
        int czz = 0;
        for (int i = 0; i + 1 < n; i++) czz += (((j >> i) ^ (j >> (i + 1))) & 1) ? -1 : 1;
        double xr = 0.0, xi = 0.0;
        for (int i = 0; i < n; i++) {
            const uint64_t p = j ^ (1ull << i);
            xr = xr + (double)re[p];
            xi = xi + (double)im[p];
        }
        lr[j] = (T)((double)czz * (double)re[j] + hx * xr);
        li[j] = (T)((double)czz * (double)im[j] + hx * xi);
    }
}

// max |a - b| over the state, into out[0]: a grid-stride maximum per thread,
// a wavefront shuffle and an LDS step per block, then one integer atomic max
// per block on the (non-negative) double bit patterns.
template <typename T>
__global__ void __launch_bounds__(kT)
k_maxdiff(const T *__restrict__ ar, const T *__restrict__ ai, const T *__restrict__ br,
          const T *__restrict__ bi, uint64_t N, unsigned long long *__restrict__ out)
{
    __shared__ double W[kT / 64];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    double m = 0.0;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += stride)
        m = fmax(m, fmax(fabs((double)ar[j] - (double)br[j]), fabs((double)ai[j] - (double)bi[j])));
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) W[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kT / 64; w++) m = fmax(m, W[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
    }
}

// Blocks of kT threads covering `units` (per_block units each), capped at the
// dispatch limit of 2^32 - 1 work-items (element-wise kernels grid-stride).
inline unsigned blocks_for(uint64_t units, uint64_t per_block)
{
    uint64_t g = (units + per_block - 1) / per_block;
    const uint64_t cap = 0xFFFFFFFFull / kT;
    if (g > cap) g = cap;
    return (unsigned)(g == 0 ? 1 : g);
}

// Blocks of the drift check: enough to fill the GPU, each striding over the state.
unsigned maxdiff_grid(uint64_t N)
{
    const uint64_t cap = (uint64_t)std::max(rt().cu_count, 1) * 8;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(cap, (N + kT - 1) / kT));
}

uint32_t kind_bit_of(int ak)
{
    switch (ak) {
    case AK_H: return TG_H;
    case AK_X: return TG_X;
    case AK_Y: return TG_Y;
    case AK_RX: return TG_RX;
    case AK_RY: return TG_RY;
    case AK_U: return TG_U;
    case AK_NEG: return TG_NEG;
    case AK_PHASE: return TG_PHASE;
    case AK_RZ: return TG_RZ;
    default: return 0;
    }
}

bool has_param(int kind)
{
    return kind == SPZ_RX || kind == SPZ_RY || kind == SPZ_RZ || kind == SPZ_P;
}

int grad_kind(int kind)
{
    switch (kind) {
    case SPZ_RX: return GK_X;
    case SPZ_RY: return GK_Y;
    case SPZ_RZ: return GK_Z;
    case SPZ_P: return GK_P;
    default: return GK_NONE;
    }
}

// The inverse of one op (gates.rs:78-92): rotations and P negate theta,
// U(theta, phi, lambda)^-1 = U(-theta, -lambda, -phi), the rest are self-inverse.
spz_op inverse_op(const spz_op &o)
{
    spz_op r = o;
    if (has_param(o.kind)) r.theta = -o.theta;
    if (o.kind == SPZ_U) {
        r.theta = -o.theta;
        r.phi = -o.lambda;
        r.lambda = -o.phi;
    }
    return r;
}

int launch_group(spz_state_s *psi, spz_state_s *lam, int nq, uint64_t nel, const PosList &pl,
                 const GroupDesc &gd, double *partials, unsigned g)
{
    hipStream_t st = rt().stream;
#define SPZ_GRP(T, M)                                                                        \
    hipLaunchKernelGGL((k_adj_group<T, M>), dim3(g), dim3(kT), 0, st, (T *)psi->re,            \
                       (T *)psi->im, (T *)lam->re, (T *)lam->im, nel, pl, gd, partials)
    if (psi->prec == SPZ_F64) {
        if (nq == 1) SPZ_GRP(double, 1);
        else if (nq == 2) SPZ_GRP(double, 2);
        else SPZ_GRP(double, 3);
    } else {
        if (nq == 1) SPZ_GRP(float, 1);
        else if (nq == 2) SPZ_GRP(float, 2);
        else SPZ_GRP(float, 3);
    }
#undef SPZ_GRP
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

// Owns a scratch state; not copyable, so it can live in a vector only
// behind a unique_ptr (a copy would free the state twice).
struct OwnedState {
    spz_state_s *s = nullptr;
    OwnedState() = default;
    OwnedState(const OwnedState &) = delete;
    OwnedState &operator=(const OwnedState &) = delete;
    ~OwnedState() { if (s) spz_state_destroy(s); }
};

int make_like(const spz_state_s *s, OwnedState &o)
{
    return spz_state_create(s->n, s->prec, &o.s);
}

// Scratch kept between calls (hipFree synchronises the device, and small
// circuits would otherwise spend most of their time allocating): four
// grow-only arrays that back the two working states (psi copy and lambda,
// viewed with the current shape) and grow-only device buffers.  One lock
// serialises the adjoint calls that share it.
struct AdjScratch {
    std::mutex mu;
    void *arr[4] = {nullptr, nullptr, nullptr, nullptr};  // work re/im, lambda re/im
    size_t arr_cap = 0;
    spz_state_s work, lam;  // views into arr
    void *grad = nullptr, *part = nullptr, *dmax = nullptr;
    size_t grad_cap = 0, part_cap = 0;

    int reserve_states(int n, int prec)
    {
        const size_t bytes = ((size_t)1 << n) * (prec == SPZ_F64 ? 8 : 4);
        if (bytes > arr_cap) {
            for (void *&a : arr)
                if (a) {
                    SPZ_HIP(hipFree(a));
                    a = nullptr;
                }
            arr_cap = 0;
            for (void *&a : arr)
                if (hipMalloc(&a, bytes) != hipSuccess) {
                    (void)hipGetLastError();
                    for (void *&b : arr)
                        if (b) {
                            (void)hipFree(b);
                            b = nullptr;
                        }
                    return fail(SPZ_ERR_OOM, "adjoint scratch allocation failed");
                }
            arr_cap = bytes;
        }
        return SPZ_OK;
    }

    int states_like(const spz_state_s *s)
    {
        SPZ_TRY(reserve_states(s->n, s->prec));
        work.n = lam.n = s->n;
        work.prec = lam.prec = s->prec;
        work.N = lam.N = s->N;
        work.re = arr[0];
        work.im = arr[1];
        lam.re = arr[2];
        lam.im = arr[3];
        return SPZ_OK;
    }

    static int grow(void **p, size_t *cap, size_t bytes)
    {
        if (bytes <= *cap) return SPZ_OK;
        if (*p) SPZ_HIP(hipFree(*p));
        *p = nullptr;
        *cap = 0;
        if (hipMalloc(p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            return fail(SPZ_ERR_OOM, "adjoint scratch allocation failed");
        }
        *cap = bytes;
        return SPZ_OK;
    }

    int buffers(size_t grad_bytes, size_t part_bytes)
    {
        SPZ_TRY(grow(&grad, &grad_cap, grad_bytes));
        SPZ_TRY(grow(&part, &part_cap, part_bytes));
        if (!dmax) {
            if (hipMalloc(&dmax, sizeof(unsigned long long)) != hipSuccess) {
                (void)hipGetLastError();
                dmax = nullptr;
                return fail(SPZ_ERR_OOM, "adjoint scratch allocation failed");
            }
        }
        return SPZ_OK;
    }
};

size_t grad_bytes(uint64_t n_ops) { return std::max<uint64_t>(n_ops, 1) * sizeof(double); }

size_t part_bytes(uint64_t N)
{
    const uint64_t per_block = (uint64_t)kT * kGPerThread;
    return ((N / 2 + per_block - 1) / per_block + 1) * 2 * kGMaxOps * sizeof(double);
}

AdjScratch &adj_scratch()
{
    static AdjScratch *a = new AdjScratch();  // lives until the process ends
    return *a;
}

}  // namespace

// Size the adjoint's scratch for circuits of up to n qubits and n_ops ops, so
// that later calls allocate (and free) nothing.
int adjoint_reserve(int n, int prec, uint64_t n_ops)
{
    AdjScratch &sc = adj_scratch();
    std::lock_guard<std::mutex> lock(sc.mu);
    SPZ_TRY(sc.reserve_states(n, prec));
    return sc.buffers(grad_bytes(n_ops), part_bytes(1ull << n));
}

// The ops a differentiated circuit may hold: unitary single-target gates H..U
// (and SWAP where allowed) with 1- and 0-controls; every qubit in range, the
// target (SWAP: both qubits) not a control, no qubit both a 1- and a
// 0-control, and at most max_controls controls.  The single-circuit and the
// batched paths share this check, so a circuit is accepted or rejected the
// same way whichever path runs it.
int validate_adjoint_op(int n, const spz_op &o, uint64_t idx, bool allow_swap, int max_controls)
{
    const bool swap = o.kind == SPZ_SWAP;
    if (o.kind < SPZ_H || o.kind > SPZ_SWAP || (swap && !allow_swap))
        return fail(SPZ_ERR_GATE, "adjoint: op " + std::to_string(idx) + " is not a unitary " +
                                      (allow_swap ? "single-target gate or SWAP"
                                                  : "single-target gate"));
    const uint64_t all = n >= 64 ? ~0ull : (1ull << n) - 1;
    const uint64_t tb = o.target >= 0 && o.target < n ? 1ull << o.target : 0;
    const uint64_t t1b = o.t1 >= 0 && o.t1 < n ? 1ull << o.t1 : 0;
    if (!tb || (swap && (!t1b || t1b == tb)) || ((o.ones | o.zeros) & ~all) ||
        (o.ones & o.zeros) || ((o.ones | o.zeros) & (tb | (swap ? t1b : 0))))
        return fail(SPZ_ERR_QUBIT, "adjoint: op " + std::to_string(idx) +
                                       " has a qubit out of range or used twice");
    if (__builtin_popcountll(o.ones | o.zeros) > max_controls)
        return fail(SPZ_ERR_GATE, "adjoint: op " + std::to_string(idx) + " has more than " +
                                      std::to_string(max_controls) + " controls");
    return SPZ_OK;
}

int adjoint_grad(spz_state_s *s, const spz_op *ops, uint64_t n_ops, double hx, int ckpt_every,
                 double *grad, double *energy, double *drift)
{
    SPZ_TRY(check_state(s));
    if (n_ops && (!ops || !grad)) return fail(SPZ_ERR_ARG, "null ops/grad");
    for (uint64_t i = 0; i < n_ops; i++) SPZ_TRY(validate_adjoint_op(s->n, ops[i], i, true, 64));
    const int n = s->n;
    // ---- forward pass with checkpoints ---------------------------------------
    std::vector<std::unique_ptr<OwnedState>> ckpts;
    std::vector<uint64_t> ckpt_at;  // checkpoint j holds psi after op ckpt_at[j]
    spz_exec_opts eo;
    memset(&eo, 0, sizeof eo);
    eo.mode = SPZ_EXEC_AUTO;
    if (ckpt_every > 0 && n_ops > 0) {
        uint64_t done = 0;
        while (done < n_ops) {
            uint64_t len = std::min<uint64_t>((uint64_t)ckpt_every, n_ops - done);
            SPZ_TRY(execute_ops(s, ops + done, len, nullptr, &eo, nullptr));
            done += len;
            if (done < n_ops) {
                ckpts.push_back(std::make_unique<OwnedState>());
                SPZ_TRY(make_like(s, *ckpts.back()));
                SPZ_TRY(spz_state_copy(ckpts.back()->s, s));
                ckpt_at.push_back(done - 1);
            }
        }
    } else if (n_ops) {
        SPZ_TRY(execute_ops(s, ops, n_ops, nullptr, &eo, nullptr));
    }
    // ---- lambda = H psi, energy --------------------------------------------
    AdjScratch &sc = adj_scratch();
    std::lock_guard<std::mutex> lock(sc.mu);
    SPZ_TRY(sc.states_like(s));
    spz_state_s *const wk = &sc.work, *const lm = &sc.lam;  // psi copy, lambda
    SPZ_TRY(spz_state_copy(wk, s));
    {
        LaunchTimer tm(SPZ_K_ADJOINT, (double)s->N * (s->prec == SPZ_F64 ? 16.0 : 8.0) * (n + 2));
        if (s->prec == SPZ_F64)
            hipLaunchKernelGGL((k_ising_apply<double>), dim3(blocks_for(s->N, kT)), dim3(kT), 0,
                               rt().stream, (const double *)s->re, (const double *)s->im,
                               (double *)lm->re, (double *)lm->im, s->N, n, hx);
        else
            hipLaunchKernelGGL((k_ising_apply<float>), dim3(blocks_for(s->N, kT)), dim3(kT), 0,
                               rt().stream, (const float *)s->re, (const float *)s->im,
                               (float *)lm->re, (float *)lm->im, s->N, n, hx);
        SPZ_HIP(hipGetLastError());
    }
    double e = 0.0;
    SPZ_TRY(reduce(s, lm, TERM_DOT_RE, 0, 0, 0.0, &e));
    if (energy) *energy = e;
    // ---- reverse sweep -------------------------------------------------------
    const uint64_t per_block = (uint64_t)kT * kGPerThread;
    SPZ_TRY(sc.buffers(grad_bytes(n_ops), part_bytes(s->N)));
    SPZ_HIP(hipMemsetAsync(sc.grad, 0, std::max<uint64_t>(n_ops, 1) * sizeof(double), rt().stream));
    SPZ_HIP(hipMemsetAsync(sc.dmax, 0, sizeof(unsigned long long), rt().stream));
    double *gdev = (double *)sc.grad, *part = (double *)sc.part;
    void *dmax_p = sc.dmax;
    const double amp = s->prec == SPZ_F64 ? 16.0 : 8.0;
    // run_lo[k]: first index of the run of parameter-free ops that ends at k
    std::vector<uint64_t> run_lo(n_ops);
    for (uint64_t k = 0; k < n_ops; k++)
        run_lo[k] = has_param(ops[k].kind) ? k + 1 : (k > 0 && !has_param(ops[k - 1].kind) ? run_lo[k - 1] : k);
    const bool tiled = n >= 13;  // the fused executor tiles 12 qubits
    constexpr uint64_t kMinRun = 4;
    int next_ck = (int)ckpt_at.size() - 1;
    uint64_t ii = n_ops;  // ops [0, ii) remain
    while (ii > 0) {
        const uint64_t hi = ii - 1;
        // a block of ops must not step over a checkpoint restore point
        const uint64_t lo_lim = next_ck >= 0 ? ckpt_at[next_ck] + 1 : 0;
        const spz_op &o = ops[hi];
        uint64_t lo = hi;
        if (tiled && !has_param(o.kind) && hi + 1 - std::max(run_lo[hi], lo_lim) >= kMinRun) {
            lo = std::max(run_lo[hi], lo_lim);
            std::vector<spz_op> inv;
            inv.reserve(hi + 1 - lo);
            for (uint64_t k = hi + 1; k-- > lo;) inv.push_back(inverse_op(ops[k]));
            SPZ_TRY(execute_ops(wk, inv.data(), inv.size(), nullptr, &eo, nullptr));
            SPZ_TRY(execute_ops(lm, inv.data(), inv.size(), nullptr, &eo, nullptr));
        } else if (o.kind == SPZ_SWAP) {  // self-inverse, no parameter
            SPZ_TRY(apply_gate(wk, SPZ_SWAP, 0, 0, 0, o.target, o.t1, o.ones, o.zeros));
            SPZ_TRY(apply_gate(lm, SPZ_SWAP, 0, 0, 0, o.target, o.t1, o.ones, o.zeros));
        } else {
            // greedy group: ops hi, hi-1, ... while the targets fit in kGMaxQ qubits
            GroupDesc gd;
            memset(&gd, 0, sizeof gd);
            GradIdx gi;
            int nq = 0, ngrad = 0;
            uint64_t qmask = 0, all_ones = ~0ull, all_zeros = ~0ull;
            for (uint64_t k = hi + 1; k-- > lo_lim;) {
                const spz_op &g = ops[k];
                if (g.kind == SPZ_SWAP || gd.nop == kGMaxOps) break;
                if (k != hi && tiled && !has_param(g.kind) &&
                    k + 1 - std::max(run_lo[k], lo_lim) >= kMinRun)
                    break;  // leave a long parameter-free run to the tile executor
                const uint64_t tb = 1ull << g.target;
                if (!(qmask & tb)) {
                    if (nq == kGMaxQ) break;
                    gd.q[nq++] = (int8_t)g.target;
                    qmask |= tb;
                }
                GroupOp &go = gd.op[gd.nop];
                double th = g.theta, ph = g.phi, la = g.lambda;
                if (has_param(g.kind)) th = -th;
                if (g.kind == SPZ_U) { th = -g.theta; ph = -g.lambda; la = -g.phi; }  // gates.rs:78-92
                Coef kc;
                SPZ_TRY(make_coef(g.kind, th, ph, la, &kc));
                go.ak = kc.ak;
                go.kb = kind_bit_of(kc.ak);
                if (!go.kb) return fail(SPZ_ERR_GATE, "adjoint: unsupported arithmetic");
                memcpy(go.c, kc.c, sizeof go.c);
                go.gk = grad_kind(g.kind);
                for (int b = 0; b < nq; b++)
                    if (gd.q[b] == g.target) go.tq = b;
                go.cm = g.ones | g.zeros;
                go.ones = g.ones;
                all_ones &= g.ones;
                all_zeros &= g.zeros;
                gi.i[gd.nop] = go.gk != GK_NONE ? (int64_t)k : -1;
                ngrad += go.gk != GK_NONE;
                gd.nop++;
                lo = k;
            }
            // common controls outside Q are fixed positions of every element
            all_ones &= ~qmask;
            all_zeros &= ~qmask;
            gd.base_ones = all_ones;
            PosList pl;
            const int np = build_poslist(qmask | all_ones | all_zeros, &pl);
            const uint64_t nel = s->N >> np;
            const unsigned g = blocks_for(nel, per_block);
            {
                LaunchTimer tm(SPZ_K_ADJOINT, 2.0 * 2.0 * (double)(nel << nq) * amp);
                SPZ_TRY(launch_group(wk, lm, nq, nel, pl, gd, part, g));
            }
            if (ngrad) {  // one block per op; ops without a gradient write nothing
                hipLaunchKernelGGL(k_adj_finish, dim3(gd.nop), dim3(kT), 0, rt().stream, part,
                                   (uint64_t)g, gi, gdev);
                SPZ_HIP(hipGetLastError());
            }
        }
        // work now holds psi_{lo-1}: restore from a checkpoint taken after op lo-1
        if (next_ck >= 0 && lo >= 1 && ckpt_at[next_ck] == lo - 1) {
            spz_state_s *ck = ckpts[next_ck]->s;
            if (s->prec == SPZ_F64)
                hipLaunchKernelGGL((k_maxdiff<double>), dim3(maxdiff_grid(s->N)), dim3(kT), 0,
                                   rt().stream, (const double *)wk->re,
                                   (const double *)wk->im, (const double *)ck->re,
                                   (const double *)ck->im, s->N, (unsigned long long *)dmax_p);
            else
                hipLaunchKernelGGL((k_maxdiff<float>), dim3(maxdiff_grid(s->N)), dim3(kT), 0,
                                   rt().stream, (const float *)wk->re,
                                   (const float *)wk->im, (const float *)ck->re,
                                   (const float *)ck->im, s->N, (unsigned long long *)dmax_p);
            SPZ_HIP(hipGetLastError());
            SPZ_TRY(spz_state_copy(wk, ck));
            next_ck--;
        }
        ii = lo;
    }
    SPZ_HIP(hipMemcpyAsync(grad, gdev, n_ops * sizeof(double), hipMemcpyDeviceToHost, rt().stream));
    unsigned long long mx = 0;
    SPZ_HIP(hipMemcpyAsync(&mx, dmax_p, sizeof mx, hipMemcpyDeviceToHost, rt().stream));
    SPZ_HIP(hipStreamSynchronize(rt().stream));
    if (drift) {
        double d;
        memcpy(&d, &mx, sizeof d);
        *drift = d;
    }
    return SPZ_OK;
}

}  // namespace spz

extern "C" int spz_adjoint_grad(spz_state s, const spz_op *ops, uint64_t n_ops, double hx,
                                int ckpt_every, double *grad, double *energy, double *drift)
{
    SPZ_API_LOCK;
    return spz::adjoint_grad(s, ops, n_ops, hx, ckpt_every, grad, energy, drift);
}
