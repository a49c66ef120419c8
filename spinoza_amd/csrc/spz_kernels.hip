// spz_kernels.hip -- plain (one gate per HBM pass) kernels for gfx950.
//
// Layout: two arrays re[N], im[N] (FP64 or FP32), N = 2^n.  A gate on target
// t pairs amplitude s0 (bit t clear) with s1 = s0 | 2^t.  The launch maps a
// flat pair index to s0 by inserting zero bits at the ascending positions of
// the target and every control, then OR-ing the controls that must read 1,
// so consecutive lanes touch consecutive addresses whenever the lowest
// inserted position is >= 1 (paired coalesced streams), and each lane then
// moves two consecutive pairs with 2-element vector loads.  Target 0 loads
// the pair itself as one 2-element vector.
//
// Every kernel is a grid-stride loop over "units" so that one launch of
// about 2048 workgroups keeps all 256 CUs busy for any state size.
#include <stdlib.h>

#include <type_traits>

#include "spz_internal.h"

namespace spz {

namespace {

constexpr int kThreads = 256;

enum Mode { MODE_VEC = 0, MODE_T0 = 1, MODE_SCALAR = 2 };

// Internal arithmetic kinds for the measurement collapse (k[0] = 1/sqrt(p)).
constexpr int AK_COLLAPSE0 = 20;   // keep |0> * scale, zero |1>
constexpr int AK_COLLAPSE1 = 21;   // keep |1> * scale, zero |0>
constexpr int AK_COLLAPSE1R = 22;  // keep |1> * scale, then X (reset)

template <int AK>
__device__ __forceinline__ void pair_op(const double *k, double &a, double &b, double &c,
                                        double &d)
{
    if constexpr (AK == AK_COLLAPSE0) {
        a = a * k[0]; b = b * k[0]; c = 0.0; d = 0.0;
    } else if constexpr (AK == AK_COLLAPSE1) {
        c = c * k[0]; d = d * k[0]; a = 0.0; b = 0.0;
    } else if constexpr (AK == AK_COLLAPSE1R) {
        a = c * k[0]; b = d * k[0]; c = 0.0; d = 0.0;
    } else {
        pair_update(AK, k, a, b, c, d);
    }
}

template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

// Exact grids (one unit per thread) stream HBM markedly faster on MI355X than
// capped grid-stride launches (measured 5.5-6.0 vs 4.8-5.0 TB/s for paired
// streams); the grid-stride loop only engages beyond 2^30 workgroups.
// One unit per thread up to the dispatch limit (a grid of at most 2^32 - 1
// work-items); beyond it the grid-stride loops of the kernels take over.
inline unsigned grid_for(uint64_t units)
{
    uint64_t g = (units + kThreads - 1) / kThreads;
    const uint64_t cap = 0xFFFFFFFFull / kThreads;
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (unsigned)g;
}

// Lane width of the plain kernels: 8-byte lanes (one pair per thread) by
// default; SPZ_PLAIN_VEC=1 selects 16-byte lanes (two pairs per thread).
inline bool plain_vec()
{
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("SPZ_PLAIN_VEC");
        v = (e && e[0] == '1') ? 1 : 0;
    }
    return v == 1;
}

// ---------------------------------------------------------------------------
// Pair kernel
// ---------------------------------------------------------------------------
template <typename T, int AK, int MODE>
__global__ void __launch_bounds__(kThreads)
k_pair(T *__restrict__ re, T *__restrict__ im, uint64_t units, int target, uint64_t ones,
       PosList pl, Coef kc)
{
    using V = typename Vec2<T>::type;
    const uint64_t tb = 1ull << target;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < units; u += stride) {
        if constexpr (MODE == MODE_VEC) {
            uint64_t s0 = insert_zeros(u << 1, pl.pos, pl.np) | ones;
            uint64_t s1 = s0 | tb;
            V r0 = *reinterpret_cast<const V *>(re + s0);
            V i0 = *reinterpret_cast<const V *>(im + s0);
            V r1 = *reinterpret_cast<const V *>(re + s1);
            V i1 = *reinterpret_cast<const V *>(im + s1);
            double a0 = r0.x, b0 = i0.x, c0 = r1.x, d0 = i1.x;
            double a1 = r0.y, b1 = i0.y, c1 = r1.y, d1 = i1.y;
            pair_op<AK>(kc.c, a0, b0, c0, d0);
            pair_op<AK>(kc.c, a1, b1, c1, d1);
            r0.x = (T)a0; r0.y = (T)a1; i0.x = (T)b0; i0.y = (T)b1;
            r1.x = (T)c0; r1.y = (T)c1; i1.x = (T)d0; i1.y = (T)d1;
            *reinterpret_cast<V *>(re + s0) = r0;
            *reinterpret_cast<V *>(im + s0) = i0;
            *reinterpret_cast<V *>(re + s1) = r1;
            *reinterpret_cast<V *>(im + s1) = i1;
        } else if constexpr (MODE == MODE_T0) {
            uint64_t s0 = insert_zeros(u, pl.pos, pl.np) | ones;
            V r = *reinterpret_cast<const V *>(re + s0);
            V i = *reinterpret_cast<const V *>(im + s0);
            double a = r.x, b = i.x, c = r.y, d = i.y;
            pair_op<AK>(kc.c, a, b, c, d);
            r.x = (T)a; r.y = (T)c; i.x = (T)b; i.y = (T)d;
            *reinterpret_cast<V *>(re + s0) = r;
            *reinterpret_cast<V *>(im + s0) = i;
        } else {
            uint64_t s0 = insert_zeros(u, pl.pos, pl.np) | ones;
            uint64_t s1 = s0 | tb;
            double a = re[s0], b = im[s0], c = re[s1], d = im[s1];
            pair_op<AK>(kc.c, a, b, c, d);
            re[s0] = (T)a; im[s0] = (T)b; re[s1] = (T)c; im[s1] = (T)d;
        }
    }
}

// ---------------------------------------------------------------------------
// Diagonal kernel: visits only amplitudes whose fixed bits match.
// ---------------------------------------------------------------------------
template <typename T, int AK, int MODE>
__global__ void __launch_bounds__(kThreads)
k_diag(T *__restrict__ re, T *__restrict__ im, uint64_t units, int target, uint64_t ones,
       PosList pl, Coef kc)
{
    using V = typename Vec2<T>::type;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < units; u += stride) {
        if constexpr (MODE == MODE_VEC) {
            uint64_t j = insert_zeros(u << 1, pl.pos, pl.np) | ones;
            V r = *reinterpret_cast<const V *>(re + j);
            V i = *reinterpret_cast<const V *>(im + j);
            double x0 = r.x, y0 = i.x, x1 = r.y, y1 = i.y;
            diag_update(AK, kc.c, (int)((j >> target) & 1), x0, y0);
            diag_update(AK, kc.c, (int)(((j + 1) >> target) & 1), x1, y1);
            r.x = (T)x0; r.y = (T)x1; i.x = (T)y0; i.y = (T)y1;
            *reinterpret_cast<V *>(re + j) = r;
            *reinterpret_cast<V *>(im + j) = i;
        } else {
            uint64_t j = insert_zeros(u, pl.pos, pl.np) | ones;
            double x = re[j], y = im[j];
            diag_update(AK, kc.c, (int)((j >> target) & 1), x, y);
            re[j] = (T)x; im[j] = (T)y;
        }
    }
}

// ---------------------------------------------------------------------------
// SWAP kernel (gates.rs:1376-1386): exchange |..0_q0..1_q1..> and |..1_q0..0_q1..>
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__global__ void __launch_bounds__(kThreads)
k_swap(T *__restrict__ re, T *__restrict__ im, uint64_t units, int q0, int q1, uint64_t ones,
       PosList pl)
{
    using V = typename Vec2<T>::type;
    const uint64_t b0 = 1ull << q0, b1 = 1ull << q1;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < units; u += stride) {
        if constexpr (MODE == MODE_VEC) {
            uint64_t base = insert_zeros(u << 1, pl.pos, pl.np) | ones;
            uint64_t sa = base | b1, sb = base | b0;
            V ra = *reinterpret_cast<const V *>(re + sa);
            V ia = *reinterpret_cast<const V *>(im + sa);
            V rb = *reinterpret_cast<const V *>(re + sb);
            V ib = *reinterpret_cast<const V *>(im + sb);
            *reinterpret_cast<V *>(re + sa) = rb;
            *reinterpret_cast<V *>(im + sa) = ib;
            *reinterpret_cast<V *>(re + sb) = ra;
            *reinterpret_cast<V *>(im + sb) = ia;
        } else {
            uint64_t base = insert_zeros(u, pl.pos, pl.np) | ones;
            uint64_t sa = base | b1, sb = base | b0;
            T r = re[sa], i = im[sa];
            re[sa] = re[sb]; im[sa] = im[sb];
            re[sb] = r; im[sb] = i;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
k_fill_basis(T *__restrict__ re, T *__restrict__ im, uint64_t N, uint64_t index)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += stride) {
        re[j] = (T)(j == index ? 1.0 : 0.0);
        im[j] = (T)0.0;
    }
}

// ---------------------------------------------------------------------------
// Dense block (transform_u): each thread owns one 2^M-amplitude vector.
// Row dot products accumulate in the order of unitary_col_vec_mul
// (gates.rs:1405-1419).
// ---------------------------------------------------------------------------
template <typename T, int M>
__global__ void __launch_bounds__(kThreads)
k_dense(T *__restrict__ re, T *__restrict__ im, uint64_t units, int t, int control,
        const double *__restrict__ ur, const double *__restrict__ ui)
{
    constexpr int D = 1 << M;
    __shared__ double sr[D * D], si[D * D];
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) { sr[i] = ur[i]; si[i] = ui[i]; }
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t lowmask = (1ull << t) - 1ull;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < units; u += stride) {
        uint64_t base = ((u >> t) << (t + M)) | (u & lowmask);
        if (control >= 0 && !((base >> control) & 1)) continue;
        double vr[D], vi[D];
#pragma unroll
        for (int j = 0; j < D; j++) {
            uint64_t kk = base + ((uint64_t)j << t);
            vr[j] = re[kk]; vi[j] = im[kk];
        }
#pragma unroll
        for (int row = 0; row < D; row++) {
            double dr = 0.0, di = 0.0;
#pragma unroll
            for (int col = 0; col < D; col++) {
                double a = sr[row * D + col], b = si[row * D + col];
                dr += a * vr[col] - b * vi[col];
                di += a * vi[col] + b * vr[col];
            }
            uint64_t kk = base + ((uint64_t)row << t);
            re[kk] = (T)dr; im[kk] = (T)di;
        }
    }
}

// ---------------------------------------------------------------------------
// Fixed-tree compensated reduction (T = 256 threads, E = 16 terms each).
// ---------------------------------------------------------------------------
constexpr int kRE = 16;
constexpr uint64_t kRB = (uint64_t)kThreads * kRE;

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e)
{
    double ss = a + b;
    double bb = ss - a;
    double aa = ss - bb;
    e = (a - aa) + (b - bb);
    s = ss;
}

template <typename T, typename T2, int TERM>
__device__ __forceinline__ double term(const T *__restrict__ re, const T *__restrict__ im,
                                       const T2 *__restrict__ re2, const T2 *__restrict__ im2,
                                       uint64_t j, int q0, int q1, int n, double hx)
{
    if constexpr (TERM == TERM_NORM) {
        double x = re[j], y = im[j];
        return x * x + y * y;
    } else if constexpr (TERM == TERM_PROB0) {
        uint64_t lo = j & ((1ull << q0) - 1ull);
        uint64_t a = ((j >> q0) << (q0 + 1)) | lo;
        double x = re[a], y = im[a];
        return x * x + y * y;
    } else if constexpr (TERM == TERM_X) {
        uint64_t p = j ^ (1ull << q0);
        double a = re[j], b = im[j], c = re[p], d = im[p];
        return a * c + b * d;
    } else if constexpr (TERM == TERM_Y) {
        uint64_t p = j ^ (1ull << q0);
        double a = re[j], b = im[j];
        double vr, vi;
        if ((j >> q0) & 1) { vr = -(double)im[p]; vi = re[p]; }
        else { vr = im[p]; vi = -(double)re[p]; }
        return a * vr + b * vi;
    } else if constexpr (TERM == TERM_Z) {
        double a = re[j], b = im[j];
        double vr = a, vi = b;
        if ((j >> q0) & 1) { vr = -vr; vi = -vi; }
        return a * vr + b * vi;
    } else if constexpr (TERM == TERM_ZZ) {
        double x = re[j], y = im[j];
        double w = x * x + y * y;
        return (((j >> q0) ^ (j >> q1)) & 1) ? -w : w;
    } else if constexpr (TERM == TERM_ISING) {
        int czz = 0;
        for (int i = 0; i + 1 < n; i++) czz += (((j >> i) ^ (j >> (i + 1))) & 1) ? -1 : 1;
        double xr = 0.0, xi = 0.0;
        for (int i = 0; i < n; i++) {
            uint64_t p = j ^ (1ull << i);
            xr = xr + (double)re[p];
            xi = xi + (double)im[p];
        }
        double a = re[j], b = im[j];
        double lr = (double)czz * a + hx * xr;
        double li = (double)czz * b + hx * xi;
        return a * lr + b * li;
    } else if constexpr (TERM == TERM_DOT_RE) {
        double a = re[j], b = im[j], c = re2[j], d = im2[j];
        return a * c + b * d;
    } else {  // TERM_DOT_IM: Im(conj(a) b) = a_re*b_im - a_im*b_re
        double a = re[j], b = im[j], c = re2[j], d = im2[j];
        return a * d - b * c;
    }
}

__device__ __forceinline__ void tree256(double *S, double *C)
{
    const int tid = threadIdx.x;
    for (int h = kThreads / 2; h >= 1; h >>= 1) {
        if (tid < h) {
            double s, e;
            two_sum(S[tid], S[tid + h], s, e);
            C[tid] = (C[tid] + C[tid + h]) + e;
            S[tid] = s;
        }
        __syncthreads();
    }
}

// T2: storage type of the second state (dot products of an FP32 and an FP64
// state are summed in FP64 like the rest).
template <typename T, typename T2, int TERM>
__global__ void __launch_bounds__(kThreads)
k_reduce1(const T *__restrict__ re, const T *__restrict__ im, const T2 *__restrict__ re2,
          const T2 *__restrict__ im2, uint64_t M, int q0, int q1, int n, double hx,
          double *__restrict__ partials)
{
    __shared__ double S[kThreads], C[kThreads];
    const int tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * kRB;
    double s = 0.0, c = 0.0;
#pragma unroll 4
    for (int e = 0; e < kRE; e++) {
        uint64_t j = base + (uint64_t)e * kThreads + tid;
        if (j < M) {
            double x = term<T, T2, TERM>(re, im, re2, im2, j, q0, q1, n, hx);
            double ns, er;
            two_sum(s, x, ns, er);
            s = ns;
            c = c + er;
        }
    }
    S[tid] = s;
    C[tid] = c;
    __syncthreads();
    tree256(S, C);
    if (tid == 0) {
        partials[2 * blockIdx.x] = S[0];
        partials[2 * blockIdx.x + 1] = C[0];
    }
}

__global__ void __launch_bounds__(kThreads)
k_reduce2(const double *__restrict__ partials, uint64_t nb, double *__restrict__ out)
{
    __shared__ double S[kThreads], C[kThreads];
    const int tid = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (uint64_t j = tid; j < nb; j += kThreads) {
        double ps = partials[2 * j], pc = partials[2 * j + 1];
        double ns, e;
        two_sum(s, ps, ns, e);
        c = (c + pc) + e;
        s = ns;
    }
    S[tid] = s;
    C[tid] = c;
    __syncthreads();
    tree256(S, C);
    if (tid == 0) out[0] = S[0] + C[0];
}

// Single-block finish when there is one level-1 block.
__global__ void k_finish1(const double *__restrict__ partials, double *__restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = partials[0] + partials[1];
}

// ---------------------------------------------------------------------------
// Inverse-CDF sampling: one thread per uniform (rule restated in the oracle).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_sample(const T *__restrict__ re, const T *__restrict__ im, uint64_t N,
         const double *__restrict__ P, const double *__restrict__ W, uint64_t nb,
         int64_t lastnz, const double *__restrict__ U, uint64_t k, uint64_t *__restrict__ out)
{
    uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= k) return;
    double total = P[nb];
    double x = U[s] * total;
    uint64_t lo = 0, hi = nb;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (P[mid + 1] > x) hi = mid; else lo = mid + 1;
    }
    uint64_t b = lo < nb ? lo : (uint64_t)lastnz;
    double acc = P[b];
    uint64_t j0 = b * kRB, j1 = j0 + kRB;
    if (j1 > N) j1 = N;
    int64_t pick = -1, lastw = -1;
    for (uint64_t j = j0; j < j1; j++) {
        double xr = re[j], xi = im[j];
        double w = xr * xr + xi * xi;
        if (w > 0.0) lastw = (int64_t)j;
        acc = acc + w;
        if (acc > x) { pick = (int64_t)j; break; }
    }
    if (pick < 0) pick = lastw;
    out[s] = (uint64_t)pick;
    (void)W;
}

// ---------------------------------------------------------------------------
// Dispatch helpers
// ---------------------------------------------------------------------------
template <typename T, int AK>
int pair_dispatch(spz_state_s *s, const Coef &k, int target, uint64_t ones, const PosList &pl,
                  uint64_t npairs, double bytes)
{
    T *re = (T *)s->re, *im = (T *)s->im;
    hipStream_t st = rt().stream;
    LaunchTimer tm(AK >= AK_COLLAPSE0 ? SPZ_K_OTHER : SPZ_K_PAIR, bytes);
    if (plain_vec() && pl.pos[0] >= 1 && npairs >= 2) {
        uint64_t units = npairs >> 1;
        hipLaunchKernelGGL((k_pair<T, AK, MODE_VEC>), dim3(grid_for(units)), dim3(kThreads), 0,
                           st, re, im, units, target, ones, pl, k);
    } else if (target == 0) {
        hipLaunchKernelGGL((k_pair<T, AK, MODE_T0>), dim3(grid_for(npairs)), dim3(kThreads), 0,
                           st, re, im, npairs, target, ones, pl, k);
    } else {
        hipLaunchKernelGGL((k_pair<T, AK, MODE_SCALAR>), dim3(grid_for(npairs)), dim3(kThreads),
                           0, st, re, im, npairs, target, ones, pl, k);
    }
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

template <typename T>
int pair_by_kind(spz_state_s *s, const Coef &k, int target, uint64_t ones, const PosList &pl,
                 uint64_t npairs, double bytes)
{
    switch (k.ak) {
    case AK_H: return pair_dispatch<T, AK_H>(s, k, target, ones, pl, npairs, bytes);
    case AK_X: return pair_dispatch<T, AK_X>(s, k, target, ones, pl, npairs, bytes);
    case AK_Y: return pair_dispatch<T, AK_Y>(s, k, target, ones, pl, npairs, bytes);
    case AK_RX: return pair_dispatch<T, AK_RX>(s, k, target, ones, pl, npairs, bytes);
    case AK_RY: return pair_dispatch<T, AK_RY>(s, k, target, ones, pl, npairs, bytes);
    case AK_U: return pair_dispatch<T, AK_U>(s, k, target, ones, pl, npairs, bytes);
    case AK_MAT: return pair_dispatch<T, AK_MAT>(s, k, target, ones, pl, npairs, bytes);
    case AK_COLLAPSE0: return pair_dispatch<T, AK_COLLAPSE0>(s, k, target, ones, pl, npairs, bytes);
    case AK_COLLAPSE1: return pair_dispatch<T, AK_COLLAPSE1>(s, k, target, ones, pl, npairs, bytes);
    case AK_COLLAPSE1R:
        return pair_dispatch<T, AK_COLLAPSE1R>(s, k, target, ones, pl, npairs, bytes);
    default: return fail(SPZ_ERR_GATE, "pair kernel: unsupported arithmetic kind");
    }
}

template <typename T, int AK>
int diag_dispatch(spz_state_s *s, const Coef &k, int target, uint64_t ones, const PosList &pl,
                  uint64_t namps, double bytes)
{
    T *re = (T *)s->re, *im = (T *)s->im;
    hipStream_t st = rt().stream;
    LaunchTimer tm(SPZ_K_DIAG, bytes);
    if (plain_vec() && (pl.np == 0 || pl.pos[0] >= 1) && namps >= 2) {
        uint64_t units = namps >> 1;
        hipLaunchKernelGGL((k_diag<T, AK, MODE_VEC>), dim3(grid_for(units)), dim3(kThreads), 0,
                           st, re, im, units, target, ones, pl, k);
    } else {
        hipLaunchKernelGGL((k_diag<T, AK, MODE_SCALAR>), dim3(grid_for(namps)), dim3(kThreads),
                           0, st, re, im, namps, target, ones, pl, k);
    }
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

template <typename T>
int diag_by_kind(spz_state_s *s, const Coef &k, int target, uint64_t ones, const PosList &pl,
                 uint64_t namps, double bytes)
{
    switch (k.ak) {
    case AK_NEG: return diag_dispatch<T, AK_NEG>(s, k, target, ones, pl, namps, bytes);
    case AK_PHASE: return diag_dispatch<T, AK_PHASE>(s, k, target, ones, pl, namps, bytes);
    case AK_RZ: return diag_dispatch<T, AK_RZ>(s, k, target, ones, pl, namps, bytes);
    default: return fail(SPZ_ERR_GATE, "diag kernel: unsupported arithmetic kind");
    }
}

inline double amp_bytes(const spz_state_s *s) { return s->prec == SPZ_F64 ? 16.0 : 8.0; }

}  // namespace

int build_poslist(uint64_t mask, PosList *pl)
{
    pl->np = 0;
    for (int q = 0; q < 64; q++)
        if ((mask >> q) & 1) pl->pos[pl->np++] = (int8_t)q;
    for (int i = pl->np; i < 64; i++) pl->pos[i] = 64;
    return pl->np;
}

int launch_pair(spz_state_s *s, const Coef &k, int target, uint64_t ones, uint64_t zeros)
{
    PosList pl;
    uint64_t fixed = ones | zeros | (1ull << target);
    int np = build_poslist(fixed, &pl);
    if (np > s->n) return fail(SPZ_ERR_QUBIT, "more fixed qubits than the state has");
    uint64_t npairs = s->N >> np;
    // read + write both amplitudes of every visited pair
    double bytes = 2.0 * 2.0 * (double)npairs * amp_bytes(s);
    if (s->prec == SPZ_F64) return pair_by_kind<double>(s, k, target, ones, pl, npairs, bytes);
    return pair_by_kind<float>(s, k, target, ones, pl, npairs, bytes);
}

int launch_diag(spz_state_s *s, const Coef &k, int target, uint64_t ones, uint64_t zeros)
{
    PosList pl;
    uint64_t fixed = ones | zeros;
    uint64_t ones2 = ones;
    if (k.ak != AK_RZ) {  // Z and P only touch amplitudes with the target bit set
        fixed |= 1ull << target;
        ones2 |= 1ull << target;
    }
    int np = build_poslist(fixed, &pl);
    if (np > s->n) return fail(SPZ_ERR_QUBIT, "more fixed qubits than the state has");
    uint64_t namps = s->N >> np;
    double bytes = 2.0 * (double)namps * amp_bytes(s);
    if (s->prec == SPZ_F64) return diag_by_kind<double>(s, k, target, ones2, pl, namps, bytes);
    return diag_by_kind<float>(s, k, target, ones2, pl, namps, bytes);
}

int launch_swap(spz_state_s *s, int q0, int q1, uint64_t ones, uint64_t zeros)
{
    PosList pl;
    uint64_t fixed = ones | zeros | (1ull << q0) | (1ull << q1);
    int np = build_poslist(fixed, &pl);
    uint64_t units = s->N >> np;  // pairs of swapped amplitudes
    double bytes = 2.0 * 2.0 * (double)units * amp_bytes(s);
    hipStream_t st = rt().stream;
    LaunchTimer tm(SPZ_K_PAIR, bytes);
    if (s->prec == SPZ_F64) {
        if (plain_vec() && pl.pos[0] >= 1 && units >= 2)
            hipLaunchKernelGGL((k_swap<double, MODE_VEC>), dim3(grid_for(units >> 1)),
                               dim3(kThreads), 0, st, (double *)s->re, (double *)s->im,
                               units >> 1, q0, q1, ones, pl);
        else
            hipLaunchKernelGGL((k_swap<double, MODE_SCALAR>), dim3(grid_for(units)),
                               dim3(kThreads), 0, st, (double *)s->re, (double *)s->im, units, q0,
                               q1, ones, pl);
    } else {
        if (plain_vec() && pl.pos[0] >= 1 && units >= 2)
            hipLaunchKernelGGL((k_swap<float, MODE_VEC>), dim3(grid_for(units >> 1)),
                               dim3(kThreads), 0, st, (float *)s->re, (float *)s->im, units >> 1,
                               q0, q1, ones, pl);
        else
            hipLaunchKernelGGL((k_swap<float, MODE_SCALAR>), dim3(grid_for(units)),
                               dim3(kThreads), 0, st, (float *)s->re, (float *)s->im, units, q0,
                               q1, ones, pl);
    }
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

int launch_collapse(spz_state_s *s, int target, int outcome, int reset, double scale)
{
    Coef k{};
    k.ak = outcome == 0 ? AK_COLLAPSE0 : (reset ? AK_COLLAPSE1R : AK_COLLAPSE1);
    k.c[0] = scale;
    return launch_pair(s, k, target, 0, 0);
}

int launch_fill_basis(spz_state_s *s, uint64_t index)
{
    hipStream_t st = rt().stream;
    LaunchTimer tm(SPZ_K_OTHER, (double)s->N * amp_bytes(s));
    if (s->prec == SPZ_F64)
        hipLaunchKernelGGL((k_fill_basis<double>), dim3(grid_for(s->N)), dim3(kThreads), 0, st,
                           (double *)s->re, (double *)s->im, s->N, index);
    else
        hipLaunchKernelGGL((k_fill_basis<float>), dim3(grid_for(s->N)), dim3(kThreads), 0, st,
                           (float *)s->re, (float *)s->im, s->N, index);
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

template <typename T>
static int dense_dispatch(spz_state_s *s, int m, const double *dur, const double *dui, int t,
                          int control)
{
    uint64_t units = s->N >> m;
    hipStream_t st = rt().stream;
    T *re = (T *)s->re, *im = (T *)s->im;
    LaunchTimer tm(SPZ_K_DENSE, 2.0 * (double)s->N * amp_bytes(s));
    switch (m) {
    case 1: hipLaunchKernelGGL((k_dense<T, 1>), dim3(grid_for(units)), dim3(kThreads), 0, st, re, im, units, t, control, dur, dui); break;
    case 2: hipLaunchKernelGGL((k_dense<T, 2>), dim3(grid_for(units)), dim3(kThreads), 0, st, re, im, units, t, control, dur, dui); break;
    case 3: hipLaunchKernelGGL((k_dense<T, 3>), dim3(grid_for(units)), dim3(kThreads), 0, st, re, im, units, t, control, dur, dui); break;
    case 4: hipLaunchKernelGGL((k_dense<T, 4>), dim3(grid_for(units)), dim3(kThreads), 0, st, re, im, units, t, control, dur, dui); break;
    case 5: hipLaunchKernelGGL((k_dense<T, 5>), dim3(grid_for(units)), dim3(kThreads), 0, st, re, im, units, t, control, dur, dui); break;
    default: return fail(SPZ_ERR_ARG, "dense block wider than 5 qubits");
    }
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

int launch_dense(spz_state_s *s, int m, const double *dur, const double *dui, int t, int control)
{
    if (s->prec == SPZ_F64) return dense_dispatch<double>(s, m, dur, dui, t, control);
    return dense_dispatch<float>(s, m, dur, dui, t, control);
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
template <typename T, typename T2 = T>
static int reduce_launch(const spz_state_s *a, const spz_state_s *b, int term, int q0, int q1,
                         double hx, uint64_t M, uint64_t nb, double *part)
{
    const T *re = (const T *)a->re, *im = (const T *)a->im;
    const T2 *re2 = b ? (const T2 *)b->re : nullptr, *im2 = b ? (const T2 *)b->im : nullptr;
    hipStream_t st = rt().stream;
    dim3 g((unsigned)nb), blk(kThreads);
    int n = a->n;
#define SPZ_RL(TM) hipLaunchKernelGGL((k_reduce1<T, T2, TM>), g, blk, 0, st, re, im, re2, im2, M, q0, q1, n, hx, part)
    if constexpr (!std::is_same<T, T2>::value) {  // mixed precision: dot products only
        if (term == TERM_DOT_RE) SPZ_RL(TERM_DOT_RE);
        else if (term == TERM_DOT_IM) SPZ_RL(TERM_DOT_IM);
        else return fail(SPZ_ERR_ARG, "mixed-precision reduction other than a dot product");
        SPZ_HIP(hipGetLastError());
        return SPZ_OK;
    }
    switch (term) {
    case TERM_NORM: SPZ_RL(TERM_NORM); break;
    case TERM_PROB0: SPZ_RL(TERM_PROB0); break;
    case TERM_X: SPZ_RL(TERM_X); break;
    case TERM_Y: SPZ_RL(TERM_Y); break;
    case TERM_Z: SPZ_RL(TERM_Z); break;
    case TERM_ZZ: SPZ_RL(TERM_ZZ); break;
    case TERM_ISING: SPZ_RL(TERM_ISING); break;
    case TERM_DOT_RE: SPZ_RL(TERM_DOT_RE); break;
    case TERM_DOT_IM: SPZ_RL(TERM_DOT_IM); break;
    default: return fail(SPZ_ERR_ARG, "unknown reduction term");
    }
#undef SPZ_RL
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

static int reduce_level1(const spz_state_s *a, const spz_state_s *b, int term, int q0, int q1,
                         double hx, uint64_t *nb_out)
{
    uint64_t M = term == TERM_PROB0 ? (a->N >> 1) : a->N;
    uint64_t nb = (M + kRB - 1) / kRB;
    if (nb == 0) nb = 1;
    SPZ_TRY(ensure_partials(nb));
    double per_term = amp_bytes(a) * (term == TERM_X || term == TERM_Y ? 2.0 : 1.0);
    if (term == TERM_DOT_RE || term == TERM_DOT_IM) per_term = amp_bytes(a) + amp_bytes(b);
    LaunchTimer tm(SPZ_K_REDUCE, (double)M * per_term);
    int rc;
    double *part = rt().d_partials;
    if (b && b->prec != a->prec)
        rc = a->prec == SPZ_F64 ? reduce_launch<double, float>(a, b, term, q0, q1, hx, M, nb, part)
                                : reduce_launch<float, double>(a, b, term, q0, q1, hx, M, nb, part);
    else
        rc = a->prec == SPZ_F64 ? reduce_launch<double>(a, b, term, q0, q1, hx, M, nb, part)
                                : reduce_launch<float>(a, b, term, q0, q1, hx, M, nb, part);
    *nb_out = nb;
    return rc;
}

int reduce(const spz_state_s *a, const spz_state_s *b, int term, int q0, int q1, double hx,
           double *out)
{
    uint64_t nb;
    SPZ_TRY(reduce_level1(a, b, term, q0, q1, hx, &nb));
    hipStream_t st = rt().stream;
    if (nb == 1)
        hipLaunchKernelGGL(k_finish1, dim3(1), dim3(1), 0, st, rt().d_partials, rt().d_result);
    else
        hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kThreads), 0, st, rt().d_partials, nb,
                           rt().d_result);
    SPZ_HIP(hipGetLastError());
    SPZ_HIP(hipMemcpyAsync(rt().h_result, rt().d_result, sizeof(double), hipMemcpyDeviceToHost, st));
    SPZ_HIP(hipStreamSynchronize(st));
    *out = rt().h_result[0];
    return SPZ_OK;
}

// ---------------------------------------------------------------------------
// Ising energy without allocation (spz_ising_energy_noalloc)
// ---------------------------------------------------------------------------
// The block partials and the result live in a file-scope device array of fixed
// size instead of the grow-only scratch that reduce() allocates: 2^18 blocks
// of 4096 terms (states up to 2^30 amplitudes, 30 qubits), 4 MiB of partial
// pairs plus the result.  Host code reaches it only through the address that
// hipGetSymbolAddress returns (the array's name is a host-side stand-in);
// that address is taken once, when the library initialises on its device.
// Every call shares the array; the library lock serialises calls within a
// process, and each call overwrites every pair it reads (no clearing needed).
constexpr uint64_t kStaticBlocks = 1ull << 18;
constexpr uint64_t kStaticResult = 2 * kStaticBlocks;  // index of the result slot
static __device__ __align__(256) double scratch[2 * kStaticBlocks + 32];
static double *g_static_scratch = nullptr;

int static_scratch_init()
{
    void *p = nullptr;
    SPZ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_static_scratch = static_cast<double *>(p);
    return SPZ_OK;
}

int ising_energy_noalloc(const spz_state_s *s, double hx, double *out)
{
    if (!g_static_scratch) return fail(SPZ_ERR_ARG, "library not initialised");
    const uint64_t M = s->N;
    const uint64_t nb = (M + kRB - 1) / kRB;
    if (nb > kStaticBlocks)
        return fail(SPZ_ERR_ARG, "state too large for spz_ising_energy_noalloc (at most 2^30 "
                                 "amplitudes; use spz_ising_energy)");
    double *part = g_static_scratch, *res = g_static_scratch + kStaticResult;
    const int rc = s->prec == SPZ_F64
                       ? reduce_launch<double>(s, nullptr, TERM_ISING, 0, 0, hx, M, nb, part)
                       : reduce_launch<float>(s, nullptr, TERM_ISING, 0, 0, hx, M, nb, part);
    SPZ_TRY(rc);
    hipStream_t st = rt().stream;
    if (nb == 1)
        hipLaunchKernelGGL(k_finish1, dim3(1), dim3(1), 0, st, part, res);
    else
        hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kThreads), 0, st, part, nb, res);
    SPZ_HIP(hipGetLastError());
    // the pinned result slot was allocated when the library initialised
    SPZ_HIP(hipMemcpyAsync(rt().h_result, res, sizeof(double), hipMemcpyDeviceToHost, st));
    SPZ_HIP(hipStreamSynchronize(st));
    *out = rt().h_result[0];
    return SPZ_OK;
}

int norm_block_partials(const spz_state_s *s, std::vector<double> &W)
{
    uint64_t nb;
    SPZ_TRY(reduce_level1(s, nullptr, TERM_NORM, 0, 0, 0.0, &nb));
    std::vector<double> pairs(2 * nb);
    SPZ_HIP(hipMemcpyAsync(pairs.data(), rt().d_partials, sizeof(double) * 2 * nb,
                           hipMemcpyDeviceToHost, rt().stream));
    SPZ_HIP(hipStreamSynchronize(rt().stream));
    W.resize(nb);
    for (uint64_t b = 0; b < nb; b++) W[b] = pairs[2 * b] + pairs[2 * b + 1];
    return SPZ_OK;
}

int launch_sample(const spz_state_s *s, const double *d_prefix, uint64_t nb, int64_t lastnz,
                  const double *d_u, uint64_t k, uint64_t *d_out)
{
    unsigned g = (unsigned)((k + kThreads - 1) / kThreads);
    LaunchTimer tm(SPZ_K_OTHER, 0.0);
    if (s->prec == SPZ_F64)
        hipLaunchKernelGGL((k_sample<double>), dim3(g), dim3(kThreads), 0, rt().stream,
                           (const double *)s->re, (const double *)s->im, s->N, d_prefix, nullptr,
                           nb, lastnz, d_u, k, d_out);
    else
        hipLaunchKernelGGL((k_sample<float>), dim3(g), dim3(kThreads), 0, rt().stream,
                           (const float *)s->re, (const float *)s->im, s->N, d_prefix, nullptr,
                           nb, lastnz, d_u, k, d_out);
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

}  // namespace spz
