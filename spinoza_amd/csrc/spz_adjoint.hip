// spz_adjoint.hip -- adjoint differentiation of E = <psi|H|psi> (no counterpart
// in the reference; built beside the gate path).
//
//   psi_k = U_k ... U_1 psi_0,  lambda_k = U_{k+1}^dag ... U_L^dag H psi_L.
//   For U_k = exp(-i theta G / 2) (RX, RY, RZ with G = X, Y, Z; controls
//   project onto the control subspace), dE/dtheta_k = Im <lambda_k| G~ |psi_k>,
//   and for P(theta) = diag(1, e^{i theta}), dE/dtheta_k = -2 Im <lambda_k|Pi_1~|psi_k>
//   (G~, Pi_1~: the operator restricted to amplitudes whose controls are met).
//
// Reverse sweep, one fused pass per gate: read the (psi, lambda) pairs of the
// gate, accumulate the gradient term (TwoSum-compensated FP64 per thread, a
// halving tree per block, one more block for the block partials, written to
// a device gradient array), apply U_k^dag to both, write both back.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "spz_internal.h"

namespace spz {

int execute_ops(spz_state_s *s, const spz_op *ops, uint64_t n_ops, spz_tracker *trk,
                const spz_exec_opts *opts, uint8_t *outcomes);

namespace {

constexpr int kT = 256;
constexpr int kPerThread = 8;  // pairs per thread per block

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

template <typename T, int AK, int GK>
__global__ void __launch_bounds__(kT)
k_adj_step(T *__restrict__ pr, T *__restrict__ pi, T *__restrict__ lr, T *__restrict__ li,
           uint64_t npairs, int target, uint64_t ones, PosList pl, Coef kc,
           double *__restrict__ partials)
{
    __shared__ double S[kT], C[kT];
    const uint64_t tb = 1ull << target;
    const uint64_t block0 = (uint64_t)blockIdx.x * kT * kPerThread;
    double s = 0.0, c = 0.0;
#pragma unroll 2
    for (int e = 0; e < kPerThread; e++) {
        const uint64_t u = block0 + (uint64_t)e * kT + threadIdx.x;
        if (u >= npairs) break;
        const uint64_t s0 = insert_zeros(u, pl.pos, pl.np) | ones;
        const uint64_t s1 = s0 | tb;
        double a = pr[s0], b = pi[s0], cc = pr[s1], d = pi[s1];
        double la = lr[s0], lb = li[s0], lc = lr[s1], ld = li[s1];
        if constexpr (GK != GK_NONE) {
            double t;
            if constexpr (GK == GK_X)  // Im <l|X|p> = Im(l0* p1) + Im(l1* p0)
                t = im_cj(la, lb, cc, d) + im_cj(lc, ld, a, b);
            else if constexpr (GK == GK_Y)  // Im <l|Y|p> = Re(l1* p0) - Re(l0* p1)
                t = re_cj(lc, ld, a, b) - re_cj(la, lb, cc, d);
            else if constexpr (GK == GK_Z)  // Im <l|Z|p> = Im(l0* p0) - Im(l1* p1)
                t = im_cj(la, lb, a, b) - im_cj(lc, ld, cc, d);
            else  // -2 Im <l|Pi_1|p>
                t = -2.0 * im_cj(lc, ld, cc, d);
            double ns, er;
            two_sum_d(s, t, ns, er);
            s = ns;
            c = c + er;
        }
        pair_update(AK, kc.c, a, b, cc, d);
        pair_update(AK, kc.c, la, lb, lc, ld);
        pr[s0] = (T)a; pi[s0] = (T)b; pr[s1] = (T)cc; pi[s1] = (T)d;
        lr[s0] = (T)la; li[s0] = (T)lb; lr[s1] = (T)lc; li[s1] = (T)ld;
    }
    if constexpr (GK != GK_NONE) {
        const int tid = threadIdx.x;
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
        if (tid == 0) {
            partials[2 * blockIdx.x] = S[0];
            partials[2 * blockIdx.x + 1] = C[0];
        }
    }
}

// One block: reduce nb block partials into out[0].
__global__ void __launch_bounds__(kT)
k_adj_finish(const double *__restrict__ partials, uint64_t nb, double *__restrict__ out)
{
    __shared__ double S[kT], C[kT];
    const int tid = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (uint64_t j = tid; j < nb; j += kT) {
        double ns, e;
        two_sum_d(s, partials[2 * j], ns, e);
        c = (c + partials[2 * j + 1]) + e;
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
    if (tid == 0) out[0] = S[0] + C[0];
}

// lambda = H psi, H = sum_{i<n-1} Z_i Z_{i+1} + hx sum_i X_i (same operation
// order as the reduction term used for the energy).
template <typename T>
__global__ void __launch_bounds__(kT)
k_ising_apply(const T *__restrict__ re, const T *__restrict__ im, T *__restrict__ lr,
              T *__restrict__ li, uint64_t N, int n, double hx)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
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

// max |a - b| over the state, into out[0] via an integer atomic max on the
// (non-negative) double bit patterns.
template <typename T>
__global__ void __launch_bounds__(kT)
k_maxdiff(const T *__restrict__ ar, const T *__restrict__ ai, const T *__restrict__ br,
          const T *__restrict__ bi, uint64_t N, unsigned long long *__restrict__ out)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (j < N) m = fmax(fabs((double)ar[j] - (double)br[j]), fabs((double)ai[j] - (double)bi[j]));
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

inline unsigned blocks_for(uint64_t units, uint64_t per_block)
{
    uint64_t g = (units + per_block - 1) / per_block;
    return (unsigned)(g == 0 ? 1 : g);
}

template <typename T, int AK>
int step_gk(spz_state_s *psi, spz_state_s *lam, uint64_t npairs, int target, uint64_t ones,
            const PosList &pl, const Coef &k, int gk, double *partials, unsigned g)
{
    T *pr = (T *)psi->re, *pi = (T *)psi->im, *lr = (T *)lam->re, *li = (T *)lam->im;
    hipStream_t st = rt().stream;
#define SPZ_ADJ(GK) hipLaunchKernelGGL((k_adj_step<T, AK, GK>), dim3(g), dim3(kT), 0, st, pr, pi, lr, li, npairs, target, ones, pl, k, partials)
    switch (gk) {
    case GK_NONE: SPZ_ADJ(GK_NONE); break;
    case GK_X: SPZ_ADJ(GK_X); break;
    case GK_Y: SPZ_ADJ(GK_Y); break;
    case GK_Z: SPZ_ADJ(GK_Z); break;
    case GK_P: SPZ_ADJ(GK_P); break;
    default: return fail(SPZ_ERR_ARG, "bad gradient kind");
    }
#undef SPZ_ADJ
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

template <typename T>
int step(spz_state_s *psi, spz_state_s *lam, uint64_t npairs, int target, uint64_t ones,
         const PosList &pl, const Coef &k, int gk, double *partials, unsigned g)
{
    switch (k.ak) {
    case AK_H: return step_gk<T, AK_H>(psi, lam, npairs, target, ones, pl, k, gk, partials, g);
    case AK_X: return step_gk<T, AK_X>(psi, lam, npairs, target, ones, pl, k, gk, partials, g);
    case AK_Y: return step_gk<T, AK_Y>(psi, lam, npairs, target, ones, pl, k, gk, partials, g);
    case AK_RX: return step_gk<T, AK_RX>(psi, lam, npairs, target, ones, pl, k, gk, partials, g);
    case AK_RY: return step_gk<T, AK_RY>(psi, lam, npairs, target, ones, pl, k, gk, partials, g);
    case AK_U: return step_gk<T, AK_U>(psi, lam, npairs, target, ones, pl, k, gk, partials, g);
    case AK_MAT: return step_gk<T, AK_MAT>(psi, lam, npairs, target, ones, pl, k, gk, partials, g);
    case AK_NEG: return step_gk<T, AK_NEG>(psi, lam, npairs, target, ones, pl, k, gk, partials, g);
    case AK_PHASE: return step_gk<T, AK_PHASE>(psi, lam, npairs, target, ones, pl, k, gk, partials, g);
    case AK_RZ: return step_gk<T, AK_RZ>(psi, lam, npairs, target, ones, pl, k, gk, partials, g);
    default: return fail(SPZ_ERR_GATE, "adjoint: unsupported arithmetic");
    }
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

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

}  // namespace

int adjoint_grad(spz_state_s *s, const spz_op *ops, uint64_t n_ops, double hx, int ckpt_every,
                 double *grad, double *energy, double *drift)
{
    SPZ_TRY(check_state(s));
    if (n_ops && (!ops || !grad)) return fail(SPZ_ERR_ARG, "null ops/grad");
    for (uint64_t i = 0; i < n_ops; i++) {
        const spz_op &o = ops[i];
        if (o.kind == SPZ_M || o.kind == SPZ_BITFLIP || o.kind == SPZ_UNITARY ||
            o.kind < SPZ_H || o.kind > SPZ_SWAP)
            return fail(SPZ_ERR_GATE, "adjoint: op " + std::to_string(i) +
                                          " is not a unitary single-target gate or SWAP");
    }
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
    OwnedState work, lam;
    SPZ_TRY(make_like(s, work));
    SPZ_TRY(make_like(s, lam));
    SPZ_TRY(spz_state_copy(work.s, s));
    {
        LaunchTimer tm(SPZ_K_ADJOINT, (double)s->N * (s->prec == SPZ_F64 ? 16.0 : 8.0) * (n + 2));
        if (s->prec == SPZ_F64)
            hipLaunchKernelGGL((k_ising_apply<double>), dim3(blocks_for(s->N, kT)), dim3(kT), 0,
                               rt().stream, (const double *)s->re, (const double *)s->im,
                               (double *)lam.s->re, (double *)lam.s->im, s->N, n, hx);
        else
            hipLaunchKernelGGL((k_ising_apply<float>), dim3(blocks_for(s->N, kT)), dim3(kT), 0,
                               rt().stream, (const float *)s->re, (const float *)s->im,
                               (float *)lam.s->re, (float *)lam.s->im, s->N, n, hx);
        SPZ_HIP(hipGetLastError());
    }
    double e = 0.0;
    SPZ_TRY(reduce(s, lam.s, TERM_DOT_RE, 0, 0, 0.0, &e));
    if (energy) *energy = e;
    // ---- reverse sweep -------------------------------------------------------
    DevBuf dgrad, dpart, dmax;
    SPZ_HIP(hipMalloc(&dgrad.p, std::max<uint64_t>(n_ops, 1) * sizeof(double)));
    SPZ_HIP(hipMemsetAsync(dgrad.p, 0, std::max<uint64_t>(n_ops, 1) * sizeof(double), rt().stream));
    const uint64_t max_blocks = (s->N / 2 + (uint64_t)kT * kPerThread - 1) / ((uint64_t)kT * kPerThread) + 1;
    SPZ_HIP(hipMalloc(&dpart.p, max_blocks * 2 * sizeof(double)));
    SPZ_HIP(hipMalloc(&dmax.p, sizeof(unsigned long long)));
    SPZ_HIP(hipMemsetAsync(dmax.p, 0, sizeof(unsigned long long), rt().stream));
    double *gdev = (double *)dgrad.p, *part = (double *)dpart.p;
    const double amp = s->prec == SPZ_F64 ? 16.0 : 8.0;
    int next_ck = (int)ckpt_at.size() - 1;
    for (uint64_t ii = n_ops; ii-- > 0;) {
        const spz_op &o = ops[ii];
        if (o.kind == SPZ_SWAP) {  // self-inverse, no parameter
            SPZ_TRY(apply_gate(work.s, SPZ_SWAP, 0, 0, 0, o.target, o.t1, o.ones, o.zeros));
            SPZ_TRY(apply_gate(lam.s, SPZ_SWAP, 0, 0, 0, o.target, o.t1, o.ones, o.zeros));
        } else {
            const int t = o.target;
            if (t < 0 || t >= n) return fail(SPZ_ERR_QUBIT, "adjoint: target out of range");
            // inverse gate (gates.rs:78-92)
            double th = o.theta, ph = o.phi, la = o.lambda;
            int kind = o.kind;
            if (kind == SPZ_P || kind == SPZ_RX || kind == SPZ_RY || kind == SPZ_RZ) th = -th;
            if (kind == SPZ_U) { th = -o.theta; ph = -o.lambda; la = -o.phi; }
            Coef k;
            SPZ_TRY(make_coef(kind, th, ph, la, &k));
            int gk = GK_NONE;
            if (kind == SPZ_RX) gk = GK_X;
            else if (kind == SPZ_RY) gk = GK_Y;
            else if (kind == SPZ_RZ) gk = GK_Z;
            else if (kind == SPZ_P) gk = GK_P;
            PosList pl;
            const uint64_t fixed = o.ones | o.zeros | (1ull << t);
            const int np = build_poslist(fixed, &pl);
            if (np > n) return fail(SPZ_ERR_QUBIT, "adjoint: too many fixed qubits");
            const uint64_t npairs = s->N >> np;
            const unsigned g = blocks_for(npairs, (uint64_t)kT * kPerThread);
            {
                LaunchTimer tm(SPZ_K_ADJOINT, 2.0 * 2.0 * 2.0 * (double)npairs * amp);
                if (s->prec == SPZ_F64)
                    SPZ_TRY(step<double>(work.s, lam.s, npairs, t, o.ones, pl, k, gk, part, g));
                else
                    SPZ_TRY(step<float>(work.s, lam.s, npairs, t, o.ones, pl, k, gk, part, g));
            }
            if (gk != GK_NONE) {
                hipLaunchKernelGGL(k_adj_finish, dim3(1), dim3(kT), 0, rt().stream, part,
                                   (uint64_t)g, gdev + ii);
                SPZ_HIP(hipGetLastError());
            }
        }
        // work now holds psi_{ii-1}: restore from a checkpoint taken after op ii-1
        if (next_ck >= 0 && ii >= 1 && ckpt_at[next_ck] == ii - 1) {
            spz_state_s *ck = ckpts[next_ck]->s;
            if (s->prec == SPZ_F64)
                hipLaunchKernelGGL((k_maxdiff<double>), dim3(blocks_for(s->N, kT)), dim3(kT), 0,
                                   rt().stream, (const double *)work.s->re,
                                   (const double *)work.s->im, (const double *)ck->re,
                                   (const double *)ck->im, s->N, (unsigned long long *)dmax.p);
            else
                hipLaunchKernelGGL((k_maxdiff<float>), dim3(blocks_for(s->N, kT)), dim3(kT), 0,
                                   rt().stream, (const float *)work.s->re,
                                   (const float *)work.s->im, (const float *)ck->re,
                                   (const float *)ck->im, s->N, (unsigned long long *)dmax.p);
            SPZ_HIP(hipGetLastError());
            SPZ_TRY(spz_state_copy(work.s, ck));
            next_ck--;
        }
    }
    SPZ_HIP(hipMemcpyAsync(grad, gdev, n_ops * sizeof(double), hipMemcpyDeviceToHost, rt().stream));
    unsigned long long mx = 0;
    SPZ_HIP(hipMemcpyAsync(&mx, dmax.p, sizeof mx, hipMemcpyDeviceToHost, rt().stream));
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
    return spz::adjoint_grad(s, ops, n_ops, hx, ckpt_every, grad, energy, drift);
}
