// spz_batch.hip -- batched adjoint gradients of many independent circuits
// (no counterpart in the reference).
//
// Circuits up to the switch point run inside a persistent kernel.  Small ones
// (n <= group_qubits, 12 by default) run whole inside one workgroup: the
// workgroup takes the next ticket from an atomic queue, keeps psi and lambda in
// its own slice of a scratch buffer, runs the forward gates, lambda = H psi,
// the energy, and the reverse sweep with the per-parameter reductions, all
// separated by workgroup barriers instead of kernel launches.  Larger ones run
// on a group of 2^(n - group_qubits) workgroups of one XCD: each member owns a
// contiguous share of every gate's pairs, and the members meet at a group
// barrier after every gate (a counter in memory; the members' stores reach the
// XCD's shared L2 before they arrive, and each member drops its CU's L1 copies
// after it leaves, so no L2 writeback is needed).  Every XCD has its own
// ticket queue, and a workgroup only takes tickets from the queue of the XCD
// it runs on, so a group's members always share an L2; tickets are handed out
// in order, so at most the group at the head of a queue is incomplete and its
// missing members are the next workgroups to finish their work.  Barrier
// waits are bounded: a member that waits too long raises an error flag and
// every workgroup leaves.  Larger circuits use the whole GPU one at a time
// (the single-circuit path).  The switch point is measured by
// spz_batch_autotune.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "spz_internal.h"

namespace spz {

int adjoint_grad(spz_state_s *s, const spz_op *ops, uint64_t n_ops, double hx, int ckpt_every,
                 double *grad, double *energy, double *drift);
int adjoint_reserve(int n, int prec, uint64_t n_ops);
int validate_adjoint_op(int n, const spz_op &o, uint64_t idx, bool allow_swap, int max_controls);

namespace {

constexpr int kW = 256;
constexpr int kMaxControls = 7;  // BOp::pos holds the target and up to 7 controls

// Compact gate record for the persistent kernel.
struct BOp {
    int32_t ak_fwd, ak_inv;
    int32_t target, gk;      // gk: 0 none, 1 X, 2 Y, 3 Z, 4 P (gradient term)
    int32_t np, pad;
    int8_t pos[8];           // ascending fixed positions (target + controls)
    uint64_t ones;
    double cf[8], ci[8];     // forward / inverse coefficients
};

struct BCirc {
    int32_t n, nops;
    uint64_t op0;            // first BOp
    uint64_t grad0;          // first gradient slot
    uint64_t region;         // group circuits: offset (doubles) of psi/lambda in the pool
    uint64_t part;           // group circuits: offset (pairs) of the partial sums
    uint32_t bar;            // group circuits: barrier counter index
    uint32_t pad;
};

// One unit of work: member `m` of the `g` workgroups that run circuit `c`.
struct BTicket {
    uint32_t c;
    uint16_t m, g;
};

constexpr unsigned kSpinLimit = 1u << 22;  // bounded barrier waits (~seconds)
constexpr int kWgPerCu = 2;                // default resident workgroups per CU

__device__ __forceinline__ void b_two_sum(double a, double b, double &s, double &e)
{
    double ss = a + b;
    double bb = ss - a;
    double aa = ss - bb;
    e = (a - aa) + (b - bb);
    s = ss;
}

// Workgroup sum of (s, c) pairs; result valid in thread 0.
__device__ double wg_sum(double s, double c, double *S, double *C)
{
    const int tid = threadIdx.x;
    __syncthreads();
    S[tid] = s;
    C[tid] = c;
    __syncthreads();
    for (int h = kW / 2; h >= 1; h >>= 1) {
        if (tid < h) {
            double ns, e;
            b_two_sum(S[tid], S[tid + h], ns, e);
            C[tid] = (C[tid] + C[tid + h]) + e;
            S[tid] = ns;
        }
        __syncthreads();
    }
    return S[0] + C[0];
}

// Workgroup (S, C) pair of a compensated sum; valid in thread 0.
__device__ void wg_pair(double s, double c, double *S, double *C, double &ps, double &pc)
{
    const int tid = threadIdx.x;
    __syncthreads();
    S[tid] = s;
    C[tid] = c;
    __syncthreads();
    for (int h = kW / 2; h >= 1; h >>= 1) {
        if (tid < h) {
            double ns, e;
            b_two_sum(S[tid], S[tid + h], ns, e);
            C[tid] = (C[tid] + C[tid + h]) + e;
            S[tid] = ns;
        }
        __syncthreads();
    }
    ps = S[0];
    pc = C[0];
}

// XCD (XCC) this wavefront runs on: HW_REG_XCC_ID, bits 3:0 (gfx940 family).
__device__ __forceinline__ unsigned xcc_id()
{
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;
}

// Group barrier: every member has made its stores visible in the XCD's L2
// (s_waitcnt vmcnt(0) in every wavefront), one thread counts the arrival and
// waits for `target` arrivals, then every member drops its CU's L1 lines so
// that its next loads see the other members' stores (the members share the
// XCD's L2, so no L2 writeback or invalidation is needed).  Returns false (and
// raises *err) when the wait runs past kSpinLimit polls or another member
// already gave up.
__device__ bool group_barrier(unsigned *ctr, unsigned target, int *err, int *flag)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        atomicAdd(ctr, 1u);
        unsigned spins = 0;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kSpinLimit ||
                __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                atomicExch(err, 1);
                ok = 0;
                break;
            }
        }
        *flag = ok;
    }
    __syncthreads();
    const bool ok = *flag != 0;
    asm volatile("buffer_inv sc0" ::: "memory");
    __syncthreads();
    return ok;
}

__device__ __forceinline__ void apply_ak(int ak, const double *k, double &a, double &b, double &c,
                                         double &d)
{
    switch (ak) {
    case AK_H: pair_update(AK_H, k, a, b, c, d); break;
    case AK_X: pair_update(AK_X, k, a, b, c, d); break;
    case AK_Y: pair_update(AK_Y, k, a, b, c, d); break;
    case AK_RX: pair_update(AK_RX, k, a, b, c, d); break;
    case AK_RY: pair_update(AK_RY, k, a, b, c, d); break;
    case AK_U: pair_update(AK_U, k, a, b, c, d); break;
    case AK_NEG: pair_update(AK_NEG, k, a, b, c, d); break;
    case AK_PHASE: pair_update(AK_PHASE, k, a, b, c, d); break;
    case AK_RZ: pair_update(AK_RZ, k, a, b, c, d); break;
    default: pair_update(AK_MAT, k, a, b, c, d); break;
    }
}

// Gradient term of op `gk` on one pair (psi: a, b, cc, d; lambda: la..ld).
__device__ __forceinline__ double grad_term(int gk, double a, double b, double cc, double d,
                                            double la, double lb, double lc, double ld)
{
    if (gk == 1) return (la * d - lb * cc) + (lc * b - ld * a);
    if (gk == 2) return (lc * a + ld * b) - (la * cc + lb * d);
    if (gk == 3) return (la * b - lb * a) - (lc * d - ld * cc);
    return -2.0 * (lc * d - ld * cc);
}

// lambda_j = (H psi)_j for the open Ising chain; returns Re conj(psi_j) lambda_j.
__device__ __forceinline__ double ising_row(const double *pr, const double *pi, double *lr,
                                            double *li, uint64_t j, int n, double hx)
{
    int czz = 0;
    for (int i = 0; i + 1 < n; i++) czz += (((j >> i) ^ (j >> (i + 1))) & 1) ? -1 : 1;
    double xr = 0.0, xi = 0.0;
    for (int i = 0; i < n; i++) {
        const uint64_t p = j ^ (1ull << i);
        xr = xr + pr[p];
        xi = xi + pi[p];
    }
    const double a = pr[j], b = pi[j];
    const double vr = (double)czz * a + hx * xr, vi = (double)czz * b + hx * xi;
    lr[j] = vr;
    li[j] = vi;
    return a * vr + b * vi;
}

// Two ops whose pairs are the same amplitudes, split the same way over the
// members: equal fixed positions (target and controls) and control values.
__device__ __forceinline__ bool same_pairs(const BOp &a, const BOp &b)
{
    if (a.np != b.np || a.ones != b.ones) return false;
    for (int i = 0; i < a.np; i++)
        if (a.pos[i] != b.pos[i]) return false;
    return true;
}

// A whole circuit in one workgroup (psi, lambda in the workgroup's slot).
__device__ void run_single(const BCirc &c, uint64_t ci, const BOp *__restrict__ ops, double hx,
                           double *pr, double *pi, double *lr, double *li, double *S, double *C,
                           double *__restrict__ energies, double *__restrict__ grads)
{
    const int tid = threadIdx.x;
    const int n = c.n;
    const uint64_t N = 1ull << n;
    for (uint64_t j = tid; j < N; j += kW) {
        pr[j] = j == 0 ? 1.0 : 0.0;
        pi[j] = 0.0;
    }
    __syncthreads();
    // forward
    for (int o = 0; o < c.nops; o++) {
        const BOp &op = ops[c.op0 + o];
        const uint64_t tb = 1ull << op.target;
        const uint64_t np = N >> op.np;
        for (uint64_t u = tid; u < np; u += kW) {
            const uint64_t s0 = insert_zeros(u, op.pos, op.np) | op.ones;
            const uint64_t s1 = s0 | tb;
            double a = pr[s0], b = pi[s0], cc = pr[s1], d = pi[s1];
            apply_ak(op.ak_fwd, op.cf, a, b, cc, d);
            pr[s0] = a; pi[s0] = b; pr[s1] = cc; pi[s1] = d;
        }
        __syncthreads();
    }
    // lambda = H psi and the energy Re <psi|lambda>
    double es = 0.0, ec = 0.0;
    for (uint64_t j = tid; j < N; j += kW) {
        double ns, er;
        b_two_sum(es, ising_row(pr, pi, lr, li, j, n, hx), ns, er);
        es = ns;
        ec = ec + er;
    }
    const double e = wg_sum(es, ec, S, C);
    if (tid == 0) energies[ci] = e;
    __syncthreads();
    // reverse sweep
    for (int o = c.nops - 1; o >= 0; o--) {
        const BOp &op = ops[c.op0 + o];
        const uint64_t tb = 1ull << op.target;
        const uint64_t np = N >> op.np;
        const int gk = op.gk;
        double gs = 0.0, gc = 0.0;
        for (uint64_t u = tid; u < np; u += kW) {
            const uint64_t s0 = insert_zeros(u, op.pos, op.np) | op.ones;
            const uint64_t s1 = s0 | tb;
            double a = pr[s0], b = pi[s0], cc = pr[s1], d = pi[s1];
            double la = lr[s0], lb = li[s0], lc = lr[s1], ld = li[s1];
            if (gk) {
                double ns, er;
                b_two_sum(gs, grad_term(gk, a, b, cc, d, la, lb, lc, ld), ns, er);
                gs = ns;
                gc = gc + er;
            }
            apply_ak(op.ak_inv, op.ci, a, b, cc, d);
            apply_ak(op.ak_inv, op.ci, la, lb, lc, ld);
            pr[s0] = a; pi[s0] = b; pr[s1] = cc; pi[s1] = d;
            lr[s0] = la; li[s0] = lb; lr[s1] = lc; li[s1] = ld;
        }
        if (gk) {
            const double g = wg_sum(gs, gc, S, C);
            if (tid == 0) grads[c.grad0 + o] = g;
        }
        __syncthreads();
    }
}

// Member m of the g workgroups running circuit c.  The member owns the m-th
// contiguous block of N/g amplitudes: it initialises its block and forms
// lambda on it, and it takes the m-th contiguous share of every gate's pairs,
// which for a gate whose qubits all lie below log2(N/g) ("local") are exactly
// the pairs inside its own block.  So the members only meet at a group
// barrier around non-local gates (before one, because it reads other blocks,
// and after it), before lambda (which reads every block), after lambda (the
// reverse sweep overwrites what other members read) and at the end; runs of
// local gates need none, and neither do two non-local gates in a row on the
// same fixed positions and control values (RY then RZ on one qubit): a member
// then reads and writes the same pairs in both, which no other member touches.  Partial sums go to `part` (one (S, C) pair per
// member for the energy and for every op) and member 0 adds them, in member
// order, after the last barrier.  Returns false when a barrier gave up.
__device__ bool run_group(const BCirc &c, uint64_t ci, unsigned m, unsigned g,
                          const BOp *__restrict__ ops, double hx, double *pool, unsigned *bars,
                          double *parts, int *err, double *S, double *C, int *flag,
                          double *__restrict__ energies, double *__restrict__ grads)
{
    const int tid = threadIdx.x;
    const int n = c.n;
    const uint64_t N = 1ull << n;
    double *pr = pool + c.region, *pi = pr + N, *lr = pi + N, *li = lr + N;
    double *part = parts + 2ull * c.part;  // [op 0..nops-1, energy][member][S, C]
    unsigned *bar = bars + c.bar;
    unsigned arrivals = 0;
    auto sync = [&]() {
        arrivals += g;
        return group_barrier(bar, arrivals, err, flag);
    };
    const uint64_t share = N / g, lo = share * m;
    const int lb = n - (31 - __builtin_clz(g));  // local qubits: below lb
    for (uint64_t j = lo + tid; j < lo + share; j += kW) {
        pr[j] = j == 0 ? 1.0 : 0.0;
        pi[j] = 0.0;
    }
    bool prev_nl = false;
    for (int o = 0; o < c.nops; o++) {
        const BOp &op = ops[c.op0 + o];
        const bool nl = op.pos[op.np - 1] >= lb;
        const bool same = o > 0 && nl && prev_nl && same_pairs(op, ops[c.op0 + o - 1]);
        if ((nl || prev_nl) && !same && !sync()) return false;
        prev_nl = nl;
        const uint64_t tb = 1ull << op.target;
        const uint64_t np = N >> op.np, ps = np / g, plo = ps * m;
        for (uint64_t u = plo + tid; u < plo + ps; u += kW) {
            const uint64_t s0 = insert_zeros(u, op.pos, op.np) | op.ones;
            const uint64_t s1 = s0 | tb;
            double a = pr[s0], b = pi[s0], cc = pr[s1], d = pi[s1];
            apply_ak(op.ak_fwd, op.cf, a, b, cc, d);
            pr[s0] = a; pi[s0] = b; pr[s1] = cc; pi[s1] = d;
        }
        __syncthreads();
    }
    if (!sync()) return false;
    {
        double es = 0.0, ec = 0.0;
        for (uint64_t j = lo + tid; j < lo + share; j += kW) {
            double ns, er;
            b_two_sum(es, ising_row(pr, pi, lr, li, j, n, hx), ns, er);
            es = ns;
            ec = ec + er;
        }
        double ps, pc;
        wg_pair(es, ec, S, C, ps, pc);
        if (tid == 0) {
            part[2 * ((uint64_t)c.nops * g + m)] = ps;
            part[2 * ((uint64_t)c.nops * g + m) + 1] = pc;
        }
    }
    if (!sync()) return false;
    prev_nl = false;
    for (int o = c.nops - 1; o >= 0; o--) {
        const BOp &op = ops[c.op0 + o];
        const bool nl = op.pos[op.np - 1] >= lb;
        const bool same = o + 1 < c.nops && nl && prev_nl && same_pairs(op, ops[c.op0 + o + 1]);
        if ((nl || prev_nl) && !same && !sync()) return false;
        prev_nl = nl;
        const uint64_t tb = 1ull << op.target;
        const uint64_t np = N >> op.np, ps = np / g, plo = ps * m;
        const int gk = op.gk;
        double gs = 0.0, gc = 0.0;
        for (uint64_t u = plo + tid; u < plo + ps; u += kW) {
            const uint64_t s0 = insert_zeros(u, op.pos, op.np) | op.ones;
            const uint64_t s1 = s0 | tb;
            double a = pr[s0], b = pi[s0], cc = pr[s1], d = pi[s1];
            double la = lr[s0], lb2 = li[s0], lc = lr[s1], ld = li[s1];
            if (gk) {
                double ns, er;
                b_two_sum(gs, grad_term(gk, a, b, cc, d, la, lb2, lc, ld), ns, er);
                gs = ns;
                gc = gc + er;
            }
            apply_ak(op.ak_inv, op.ci, a, b, cc, d);
            apply_ak(op.ak_inv, op.ci, la, lb2, lc, ld);
            pr[s0] = a; pi[s0] = b; pr[s1] = cc; pi[s1] = d;
            lr[s0] = la; li[s0] = lb2; lr[s1] = lc; li[s1] = ld;
        }
        if (gk) {
            double ps2, pc2;
            wg_pair(gs, gc, S, C, ps2, pc2);
            if (tid == 0) {
                part[2 * ((uint64_t)o * g + m)] = ps2;
                part[2 * ((uint64_t)o * g + m) + 1] = pc2;
            }
        }
        __syncthreads();
    }
    if (!sync()) return false;
    if (m == 0) {
        // energy (slot nops) and every parameter's gradient, members in order
        for (int o = tid; o <= c.nops; o += kW) {
            if (o < c.nops && !ops[c.op0 + o].gk) continue;
            double ts = 0.0, tc = 0.0;
            for (unsigned k = 0; k < g; k++) {
                double ns, er;
                b_two_sum(ts, part[2 * ((uint64_t)o * g + k)], ns, er);
                tc = (tc + part[2 * ((uint64_t)o * g + k) + 1]) + er;
                ts = ns;
            }
            if (o == c.nops) energies[ci] = ts + tc;
            else grads[c.grad0 + o] = ts + tc;
        }
    }
    return true;
}

// Persistent kernel: every workgroup serves the ticket queue of its own XCD
// (queue q = xcc % nq) until the queue is empty or a group barrier failed.
__global__ void __launch_bounds__(kW)
k_batch_adjoint(const BCirc *__restrict__ circs, const BTicket *__restrict__ tickets,
                const uint32_t *__restrict__ qstart, const uint32_t *__restrict__ qlen, uint32_t nq,
                const BOp *__restrict__ ops, double hx, double *__restrict__ scratch,
                uint64_t slot_amps, double *__restrict__ pool, double *__restrict__ parts,
                unsigned int *__restrict__ counters, int *__restrict__ err,
                double *__restrict__ energies, double *__restrict__ grads,
                unsigned long long *__restrict__ ticks)
{
    __shared__ double S[kW], C[kW];
    __shared__ unsigned int cur;
    __shared__ int flag;
    const int tid = threadIdx.x;
    double *pr = scratch + (uint64_t)blockIdx.x * 4 * slot_amps;
    double *pi = pr + slot_amps, *lr = pi + slot_amps, *li = lr + slot_amps;
    const uint32_t q = nq > 1 ? xcc_id() % nq : 0u;
    unsigned int *qctr = counters + q;
    unsigned int *bars = counters + nq;
    for (;;) {
        if (tid == 0) cur = atomicAdd(qctr, 1u);
        __syncthreads();
        const uint32_t t = cur;
        __syncthreads();
        if (t >= qlen[q]) break;  // uniform: every wavefront leaves together
        const BTicket tk = tickets[qstart[q] + t];
        const unsigned long long t_start = wall_clock64();
        const BCirc c = circs[tk.c];
        if (tk.g <= 1) {
            run_single(c, tk.c, ops, hx, pr, pi, lr, li, S, C, energies, grads);
        } else if (!run_group(c, tk.c, tk.m, tk.g, ops, hx, pool, bars, parts, err, S, C, &flag,
                              energies, grads)) {
            break;  // a group barrier gave up: leave (uniform, from the barrier's flag)
        }
        if (tid == 0 && ticks && tk.m == 0) {
            ticks[2 * tk.c] = t_start;
            ticks[2 * tk.c + 1] = wall_clock64();
        }
    }
}

// XCD of each of `count` workgroups (probe of the device's XCD layout).
__global__ void k_xcc_probe(unsigned int *out)
{
    if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

// Zero `count` counters (inside a captured graph, where a memset node is
// avoided) and the error flag.
__global__ void k_zero_u32(unsigned int *p, uint32_t count, int *err)
{
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0u;
    if (threadIdx.x == 0) *err = 0;
}

int lower_bop(const spz_op &o, BOp *b)
{
    memset(b, 0, sizeof *b);
    if (o.kind < SPZ_H || o.kind > SPZ_U) return fail(SPZ_ERR_GATE, "batched adjoint: unsupported gate");
    Coef kf, ki;
    SPZ_TRY(make_coef(o.kind, o.theta, o.phi, o.lambda, &kf));
    double th = o.theta, ph = o.phi, la = o.lambda;
    if (o.kind == SPZ_P || o.kind == SPZ_RX || o.kind == SPZ_RY || o.kind == SPZ_RZ) th = -th;
    if (o.kind == SPZ_U) { th = -o.theta; ph = -o.lambda; la = -o.phi; }
    SPZ_TRY(make_coef(o.kind, th, ph, la, &ki));
    b->ak_fwd = kf.ak;
    b->ak_inv = ki.ak;
    for (int j = 0; j < 8; j++) { b->cf[j] = kf.c[j]; b->ci[j] = ki.c[j]; }
    b->target = o.target;
    b->gk = o.kind == SPZ_RX ? 1 : o.kind == SPZ_RY ? 2 : o.kind == SPZ_RZ ? 3 : o.kind == SPZ_P ? 4 : 0;
    const uint64_t fixed = o.ones | o.zeros | (1ull << o.target);
    int np = 0;
    for (int q = 0; q < 64; q++)
        if ((fixed >> q) & 1) {
            if (np == 8) return fail(SPZ_ERR_GATE, "batched adjoint: more than 7 controls");
            b->pos[np++] = (int8_t)q;
        }
    b->np = np;
    b->ones = o.ones;
    return SPZ_OK;
}

int g_wg_max_qubits = -1;  // autotuned switch point (-1: not tuned yet)

// A second stream (default priority) for the persistent small-circuit kernel,
// so that it runs beside the single-circuit path of the larger circuits.
hipStream_t side_stream()
{
    static hipStream_t st = nullptr;
    if (!st && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        st = nullptr;
    }
    return st;
}

// Buffers of the persistent launch, kept across calls and grown when a call
// needs more (a free synchronises the device), and the captured launch, which
// is launched again while its arguments stay the same (a call with the same
// batch shape and the same buffers).
struct BatchBuf {
    void *p = nullptr;
    size_t cap = 0;
};
struct BatchCache {
    BatchBuf circ, ops, scr, cnt, err, tk, q, pool, parts, e, g, t;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> key;  // the captured launch's arguments
};
BatchCache g_bc;

bool grow(BatchBuf &b, size_t bytes)
{
    bytes = std::max<size_t>(bytes, 256);
    if (bytes <= b.cap) return true;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    b.cap = bytes;
    return true;
}

// Number of XCDs, from the XCC ids that workgroups report (0 when the ids do
// not cover 0..k-1: grouped circuits are then disabled).  Measured once.
int xcd_count()
{
    static int nx = -1;
    if (nx >= 0) return nx;
    nx = 0;
    const unsigned P = (unsigned)std::max(rt().cu_count, 1) * 8u;
    unsigned int *d = nullptr;
    if (hipMalloc(&d, P * sizeof(unsigned int)) != hipSuccess) {
        (void)hipGetLastError();
        return nx;
    }
    std::vector<unsigned int> h(P, 0);
    hipLaunchKernelGGL(k_xcc_probe, dim3(P), dim3(64), 0, rt().stream, d);
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpyAsync(h.data(), d, P * sizeof(unsigned int), hipMemcpyDeviceToHost,
                                   rt().stream) == hipSuccess &&
                    hipStreamSynchronize(rt().stream) == hipSuccess;
    (void)hipFree(d);
    if (!ok) return nx;
    unsigned seen = 0, mx = 0;
    for (unsigned v : h) {
        if (v >= 16) return nx;
        seen |= 1u << v;
        mx = std::max(mx, v);
    }
    if (seen == (2u << mx) - 1) nx = (int)mx + 1;
    return nx;
}

}  // namespace

}  // namespace spz

using namespace spz;

extern "C" {

int spz_batch_adjoint(const spz_batch_circuit *circuits, uint64_t count, const spz_op *ops,
                      double hx, double *energies, double *grads, const spz_batch_opts *opts,
                      spz_batch_stats *stats)
{
    SPZ_API_LOCK;
    SPZ_TRY(ensure_init());
    if (count == 0) return SPZ_OK;
    if (!circuits || !ops || !energies || !grads) return fail(SPZ_ERR_ARG, "null argument");
    int wg_max = opts && opts->wg_max_qubits > 0 ? opts->wg_max_qubits
                                                 : (g_wg_max_qubits > 0 ? g_wg_max_qubits : 14);
    const int use_graph = opts ? opts->use_graph : 1;
    // circuits above gq qubits run on groups of workgroups (gq >= 8 keeps
    // every member's share of every gate's pairs non-empty)
    const int gq_opt = opts ? opts->group_qubits : 0;
    const int gq = gq_opt < 0 ? 64 : (gq_opt == 0 ? 12 : std::max(gq_opt, 8));
    std::vector<uint64_t> small, large;
    uint64_t total_ops = 0;
    for (uint64_t i = 0; i < count; i++) {
        const spz_batch_circuit &c = circuits[i];
        if (c.n < 1 || c.n > 30 || c.n_ops < 0) return fail(SPZ_ERR_ARG, "bad circuit entry");
        // one rule for both paths (the switch point must not decide whether
        // a circuit is accepted): gates H..U, at most kMaxControls controls
        for (int o = 0; o < c.n_ops; o++) {
            const int rc = validate_adjoint_op(c.n, ops[c.op_offset + o], (uint64_t)o, false,
                                               kMaxControls);
            if (rc != SPZ_OK) {
                set_error("circuit " + std::to_string(i) + ": " + spz_last_error());
                return rc;
            }
        }
        (c.n <= wg_max ? small : large).push_back(i);
        total_ops += (uint64_t)c.n_ops;
    }
    (void)total_ops;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<double> lat(count, 0.0);
    // ---- routing: circuits above group_qubits run on 2^(n - group_qubits)
    // workgroups of one XCD; a group may take up to all of an XCD's resident
    // workgroups (larger ones take the single-circuit path).  The grid never
    // exceeds what is resident at once, and tickets are handed out in order,
    // so a group's missing members are workgroups still finishing earlier
    // tickets, none of which waits on a later one: every group fills.
    int occ = 0;
    SPZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_batch_adjoint, kW, 0));
    // resident workgroups per CU (default 2): with fewer circuits sharing a CU a
    // group's members keep pace with each other (1024 circuits, n 10..20, with
    // groups up to a whole XCD: p99 87 -> 16 ms, 214k -> 253k grads/s)
    const int cap = opts && opts->wg_per_cu != 0 ? opts->wg_per_cu : kWgPerCu;
    if (cap > 0) occ = std::min(occ, cap);
    const uint64_t grid = (uint64_t)rt().cu_count * std::max(occ, 1);
    const int nx = (gq < 64 && !small.empty()) ? xcd_count() : 0;
    const uint32_t nq = nx > 0 ? (uint32_t)nx : 1u;
    const uint64_t per_q = grid / nq;
    auto members = [&](int n) -> uint64_t { return (nx > 0 && n > gq) ? (1ull << (n - gq)) : 1ull; };
    {
        std::vector<uint64_t> keep;
        for (uint64_t i : small) {
            if (members(circuits[i].n) > 1 && members(circuits[i].n) > per_q) large.push_back(i);
            else keep.push_back(i);
        }
        small.swap(keep);
    }
    // ---- larger circuits: pre-size everything they need, so that no
    // allocation or free (which would synchronise the device) happens while
    // the small-circuit kernel runs beside them
    std::stable_sort(large.begin(), large.end(),
                     [&](uint64_t x, uint64_t y) { return circuits[x].n < circuits[y].n; });
    std::vector<spz_state> st_by_n(64, nullptr);
    auto release = [&]() {
        for (spz_state &x : st_by_n)
            if (x) spz_state_destroy(x);
    };
    if (!large.empty()) {
        int nmax = 0;
        uint64_t opsmax = 0;
        for (uint64_t i : large) {
            nmax = std::max(nmax, circuits[i].n);
            opsmax = std::max<uint64_t>(opsmax, (uint64_t)circuits[i].n_ops);
        }
        int rc = adjoint_reserve(nmax, SPZ_F64, opsmax);
        for (uint64_t i : large)
            if (rc == SPZ_OK && !st_by_n[circuits[i].n])
                rc = spz_state_create(circuits[i].n, SPZ_F64, &st_by_n[circuits[i].n]);
        if (rc != SPZ_OK) {
            release();
            return rc;
        }
    }
    // ---- circuits up to the switch point: one persistent launch on the side stream
    BatchCache &bcache = g_bc;
    void *d_circ = nullptr, *d_ops = nullptr, *d_scr = nullptr, *d_cnt = nullptr, *d_e = nullptr,
         *d_g = nullptr, *d_t = nullptr, *d_tk = nullptr, *d_q = nullptr, *d_pool = nullptr,
         *d_parts = nullptr, *d_err = nullptr;
    uint64_t gmax = 0;
    hipStream_t st = side_stream() ? side_stream() : rt().stream;
    int n_group = 0;
    if (!small.empty()) {
        // largest first: the long circuits start at once and the short ones
        // fill the workgroups that finish early
        std::stable_sort(small.begin(), small.end(), [&](uint64_t x, uint64_t y) {
            const double cx = (double)circuits[x].n_ops * (double)(1ull << circuits[x].n);
            const double cy = (double)circuits[y].n_ops * (double)(1ull << circuits[y].n);
            return cx > cy;
        });
        std::vector<BCirc> bc(small.size());
        std::vector<BOp> bo;
        int nmax_single = 1;
        uint64_t pool = 0, parts = 0;
        uint32_t nbars = 0;
        // queues: each circuit to the least-loaded queue, in cost order
        std::vector<std::vector<BTicket>> qt(nq);
        std::vector<double> qload(nq, 0.0);
        for (size_t k = 0; k < small.size(); k++) {
            const spz_batch_circuit &c = circuits[small[k]];
            bc[k].n = c.n;
            bc[k].nops = c.n_ops;
            bc[k].op0 = bo.size();
            bc[k].grad0 = c.grad_offset;
            const uint64_t g = members(c.n);
            if (g > 1) {
                bc[k].region = pool;
                bc[k].bar = nbars++;
                bc[k].part = parts;
                pool += 4ull << c.n;
                parts += ((uint64_t)c.n_ops + 1) * g;
                n_group++;
            } else {
                nmax_single = std::max(nmax_single, c.n);
            }
            for (int o = 0; o < c.n_ops; o++) {
                BOp b;
                int rc = lower_bop(ops[c.op_offset + o], &b);
                if (rc != SPZ_OK) {
                    release();
                    return rc;
                }
                bo.push_back(b);
            }
            const uint32_t q = (uint32_t)(std::min_element(qload.begin(), qload.end()) - qload.begin());
            qload[q] += (double)std::max(c.n_ops, 1) * (double)(1ull << c.n);
            for (uint64_t m = 0; m < g; m++)
                qt[q].push_back(BTicket{(uint32_t)k, (uint16_t)m, (uint16_t)g});
        }
        std::vector<BTicket> tk;
        std::vector<uint32_t> qs(2 * nq);
        for (uint32_t q = 0; q < nq; q++) {
            qs[q] = (uint32_t)tk.size();
            qs[nq + q] = (uint32_t)qt[q].size();
            tk.insert(tk.end(), qt[q].begin(), qt[q].end());
        }
        const uint64_t slot = 1ull << nmax_single;
        const uint64_t grid_used = std::min<uint64_t>(grid, nq == 1 ? tk.size() : grid);
        // scratch: 4 arrays of 2^nmax doubles per workgroup for the one-workgroup circuits
        const uint64_t scratch_bytes = grid_used * 4 * slot * sizeof(double);
        for (uint64_t i : small) gmax = std::max<uint64_t>(gmax, circuits[i].grad_offset + circuits[i].n_ops);
        const uint32_t ncnt = nq + nbars;
        BatchCache &B = bcache;
        if (!grow(B.circ, bc.size() * sizeof(BCirc)) || !grow(B.ops, bo.size() * sizeof(BOp)) ||
            !grow(B.scr, scratch_bytes) || !grow(B.cnt, ncnt * sizeof(unsigned int)) ||
            !grow(B.err, sizeof(int)) || !grow(B.tk, tk.size() * sizeof(BTicket)) ||
            !grow(B.q, qs.size() * sizeof(uint32_t)) || !grow(B.pool, pool * sizeof(double)) ||
            !grow(B.parts, 2 * parts * sizeof(double)) || !grow(B.e, small.size() * sizeof(double)) ||
            !grow(B.g, gmax * sizeof(double)) ||
            !grow(B.t, small.size() * 2 * sizeof(unsigned long long))) {
            release();
            return fail(SPZ_ERR_OOM, "batched adjoint scratch allocation failed");
        }
        d_circ = B.circ.p; d_ops = B.ops.p; d_scr = B.scr.p; d_cnt = B.cnt.p; d_err = B.err.p;
        d_tk = B.tk.p; d_q = B.q.p; d_pool = B.pool.p; d_parts = B.parts.p; d_e = B.e.p;
        d_g = B.g.p; d_t = B.t.p;
        // the gradient slots of parameter-free ops stay zero
        std::vector<double> zeros(std::max<uint64_t>(gmax, 1), 0.0);
        if (hipMemcpy(d_circ, bc.data(), bc.size() * sizeof(BCirc), hipMemcpyHostToDevice) != hipSuccess ||
            (bo.size() && hipMemcpy(d_ops, bo.data(), bo.size() * sizeof(BOp), hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(d_tk, tk.data(), tk.size() * sizeof(BTicket), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_q, qs.data(), qs.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_g, zeros.data(), zeros.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            release();
            return fail(SPZ_ERR_HIP, "batched adjoint upload failed");
        }
        auto launch = [&]() {
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, st, (unsigned int *)d_cnt, ncnt,
                               (int *)d_err);
            hipLaunchKernelGGL(k_batch_adjoint, dim3((unsigned)grid_used), dim3(kW), 0, st,
                               (const BCirc *)d_circ, (const BTicket *)d_tk,
                               (const uint32_t *)d_q, (const uint32_t *)d_q + nq, nq,
                               (const BOp *)d_ops, hx, (double *)d_scr, slot, (double *)d_pool,
                               (double *)d_parts, (unsigned int *)d_cnt, (int *)d_err,
                               (double *)d_e, (double *)d_g, (unsigned long long *)d_t);
        };
        bool ok = true;
        if (use_graph) {
            // the captured launch is reused while every argument is the same
            uint64_t hxb;
            memcpy(&hxb, &hx, sizeof hxb);
            std::vector<uint64_t> key = {grid_used, nq, ncnt, slot, hxb,
                                         (uint64_t)(uintptr_t)d_circ, (uint64_t)(uintptr_t)d_ops,
                                         (uint64_t)(uintptr_t)d_scr, (uint64_t)(uintptr_t)d_cnt,
                                         (uint64_t)(uintptr_t)d_err, (uint64_t)(uintptr_t)d_tk,
                                         (uint64_t)(uintptr_t)d_q, (uint64_t)(uintptr_t)d_pool,
                                         (uint64_t)(uintptr_t)d_parts, (uint64_t)(uintptr_t)d_e,
                                         (uint64_t)(uintptr_t)d_g, (uint64_t)(uintptr_t)d_t};
            if (!(B.exec && B.key == key)) {
                if (B.exec) (void)hipGraphExecDestroy(B.exec);
                if (B.graph) (void)hipGraphDestroy(B.graph);
                B.exec = nullptr;
                B.graph = nullptr;
                B.key.clear();
                ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
                if (ok) {
                    launch();
                    ok = hipStreamEndCapture(st, &B.graph) == hipSuccess;
                }
                if (ok) ok = hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0) == hipSuccess;
                if (ok) B.key = key;
            }
            if (ok) ok = hipGraphLaunch(B.exec, st) == hipSuccess;
        } else {
            launch();
            ok = hipGetLastError() == hipSuccess;
        }
        if (!ok) {
            (void)hipGetLastError();
            (void)hipStreamSynchronize(st);
            if (B.exec) (void)hipGraphExecDestroy(B.exec);
            if (B.graph) (void)hipGraphDestroy(B.graph);
            B.exec = nullptr;
            B.graph = nullptr;
            B.key.clear();
            release();
            return fail(SPZ_ERR_HIP, "batched adjoint launch failed");
        }
    }
    // ---- larger circuits: the single-circuit path on the library stream -----------
    int rc_large = SPZ_OK;
    for (uint64_t i : large) {
        const spz_batch_circuit &c = circuits[i];
        spz_state s = st_by_n[c.n];
        int rc = spz_state_set_basis(s, 0);
        auto a = std::chrono::steady_clock::now();
        if (rc == SPZ_OK)
            rc = adjoint_grad(s, ops + c.op_offset, (uint64_t)c.n_ops, hx, 0, grads + c.grad_offset,
                              &energies[i], nullptr);
        auto b = std::chrono::steady_clock::now();
        if (rc != SPZ_OK) {
            rc_large = rc;
            break;
        }
        lat[i] = std::chrono::duration<double, std::milli>(b - a).count();
    }
    // ---- collect the small circuits ---------------------------------------------
    int rc_small = SPZ_OK;
    if (!small.empty()) {
        if (hipStreamSynchronize(st) != hipSuccess) rc_small = fail(SPZ_ERR_HIP, "batched adjoint kernel failed");
        int kerr = 0;
        if (rc_small == SPZ_OK &&
            (hipMemcpy(&kerr, d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess || kerr))
            rc_small = fail(SPZ_ERR_HIP, "batched adjoint: a workgroup group barrier timed out");
        if (rc_small == SPZ_OK) {
            std::vector<double> e(small.size()), g(std::max<uint64_t>(gmax, 1));
            std::vector<unsigned long long> t(2 * small.size());
            if (hipMemcpy(e.data(), d_e, e.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(g.data(), d_g, g.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(t.data(), d_t, t.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess)
                rc_small = fail(SPZ_ERR_HIP, "batched adjoint download failed");
            int clk_khz = 100000;  // wall_clock64 runs at 100 MHz on gfx9
            (void)hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, rt().device);
            for (size_t k = 0; rc_small == SPZ_OK && k < small.size(); k++) {
                const uint64_t i = small[k];
                energies[i] = e[k];
                const spz_batch_circuit &c = circuits[i];
                for (int o = 0; o < c.n_ops; o++) grads[c.grad_offset + o] = g[c.grad_offset + o];
                lat[i] = (double)(t[2 * k + 1] - t[2 * k]) / (double)clk_khz;  // ms
            }
        }
    }
    release();
    if (rc_large != SPZ_OK) return rc_large;
    if (rc_small != SPZ_OK) return rc_small;
    auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        std::vector<double> l = lat;
        std::sort(l.begin(), l.end());
        stats->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->p50_ms = l[(size_t)(0.50 * (l.size() - 1))];
        stats->p99_ms = l[(size_t)(0.99 * (l.size() - 1))];
        stats->n_small = small.size();
        stats->n_large = large.size();
        stats->wg_max_qubits = wg_max;
        stats->n_group = n_group;
    }
    return SPZ_OK;
}

int spz_batch_autotune(const spz_batch_circuit *circuits, uint64_t count, const spz_op *ops,
                       int *wg_max_qubits)
{
    // For every qubit count present: run up to 8 of its circuits concurrently
    // in the persistent kernel (per-circuit latency L_wg) and one at a time on
    // the single-circuit path (T_single each).  With m circuits of that size
    // in the batch and `slots` resident workgroups, the persistent kernel needs
    // about ceil(m / slots) * L_wg, the single path m * T_single.  The switch
    // point is the largest n for which the persistent kernel wins.
    SPZ_API_LOCK;
    SPZ_TRY(ensure_init());
    if (!circuits || !ops) return fail(SPZ_ERR_ARG, "null argument");
    int occ = 0;
    SPZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_batch_adjoint, kW, 0));
    occ = std::min(occ, kWgPerCu);  // the trial batches run with the default
    const double slots = (double)rt().cu_count * std::max(occ, 1);
    std::vector<int> ns;
    for (uint64_t i = 0; i < count; i++) ns.push_back(circuits[i].n);
    std::sort(ns.begin(), ns.end());
    ns.erase(std::unique(ns.begin(), ns.end()), ns.end());
    int best = 0;
    for (int n : ns) {
        std::vector<spz_batch_circuit> pick;
        double m = 0;
        for (uint64_t i = 0; i < count; i++)
            if (circuits[i].n == n) {
                m += 1;
                if (pick.size() < 8) pick.push_back(circuits[i]);
            }
        uint64_t gsz = 0;
        for (auto &c : pick) gsz = std::max<uint64_t>(gsz, c.grad_offset + c.n_ops);
        std::vector<double> e(std::max<uint64_t>(count, 1)), g(std::max<uint64_t>(gsz, 1));
        spz_batch_opts o;
        memset(&o, 0, sizeof o);
        o.use_graph = 1;
        spz_batch_stats sw, sl;
        o.wg_max_qubits = 64;  // all in the persistent kernel
        SPZ_TRY(spz_batch_adjoint(pick.data(), pick.size(), ops, 1.0, e.data(), g.data(), &o, &sw));
        o.wg_max_qubits = n - 1;  // all on the single-circuit path
        SPZ_TRY(spz_batch_adjoint(pick.data(), pick.size(), ops, 1.0, e.data(), g.data(), &o, &sl));
        const double t_wg = std::ceil(m / slots) * sw.p99_ms;
        const double t_single = m * (sl.total_ms / (double)pick.size());
        if (t_wg <= t_single) best = n;
    }
    g_wg_max_qubits = best > 0 ? best : 1;
    if (wg_max_qubits) *wg_max_qubits = g_wg_max_qubits;
    return SPZ_OK;
}

}  // extern "C"
