// spz_batch.hip -- batched adjoint gradients of many independent circuits
// (no counterpart in the reference).
//
// Small circuits run whole inside one workgroup of a persistent kernel: the
// workgroup takes the next circuit from an atomic queue, keeps psi and lambda
// in its own slice of a scratch buffer (L2/MALL resident for n <= 16), runs
// the forward gates, lambda = H psi, the energy, and the reverse sweep with
// the per-parameter reductions, all separated by workgroup barriers instead of
// kernel launches.  Every wavefront leaves the queue loop when the counter
// passes the circuit count.  Large circuits use the whole GPU one at a time
// (the single-circuit path).  The qubit count at which to switch is measured
// by spz_batch_autotune.
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
};

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

__global__ void __launch_bounds__(kW)
k_batch_adjoint(const BCirc *__restrict__ circs, uint64_t ncirc, const BOp *__restrict__ ops,
                double hx, double *__restrict__ scratch, uint64_t slot_amps,
                unsigned int *__restrict__ counter, double *__restrict__ energies,
                double *__restrict__ grads, unsigned long long *__restrict__ ticks)
{
    __shared__ double S[kW], C[kW];
    __shared__ unsigned int cur;
    const int tid = threadIdx.x;
    double *pr = scratch + (uint64_t)blockIdx.x * 4 * slot_amps;
    double *pi = pr + slot_amps, *lr = pi + slot_amps, *li = lr + slot_amps;
    for (;;) {
        if (tid == 0) cur = atomicAdd(counter, 1u);
        __syncthreads();
        const uint64_t ci = cur;
        __syncthreads();
        if (ci >= ncirc) break;  // uniform: every wavefront leaves together
        const unsigned long long t_start = wall_clock64();
        const BCirc c = circs[ci];
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
            double ns, er;
            b_two_sum(es, a * vr + b * vi, ns, er);
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
                    double t;
                    if (gk == 1) t = (la * d - lb * cc) + (lc * b - ld * a);
                    else if (gk == 2) t = (lc * a + ld * b) - (la * cc + lb * d);
                    else if (gk == 3) t = (la * b - lb * a) - (lc * d - ld * cc);
                    else t = -2.0 * (lc * d - ld * cc);
                    double ns, er;
                    b_two_sum(gs, t, ns, er);
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
        if (tid == 0 && ticks) {
            ticks[2 * ci] = t_start;
            ticks[2 * ci + 1] = wall_clock64();
        }
    }
}

__global__ void k_zero_u32(unsigned int *p) { if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u; }

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
    // ---- small circuits: one persistent launch on the side stream ----------------
    void *d_circ = nullptr, *d_ops = nullptr, *d_scr = nullptr, *d_cnt = nullptr,
         *d_e = nullptr, *d_g = nullptr, *d_t = nullptr;
    auto cleanup = [&]() {
        for (void *p : {d_circ, d_ops, d_scr, d_cnt, d_e, d_g, d_t})
            if (p) (void)hipFree(p);
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t gmax = 0;
    hipStream_t st = side_stream() ? side_stream() : rt().stream;
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
        int nmax = 1;
        for (size_t k = 0; k < small.size(); k++) {
            const spz_batch_circuit &c = circuits[small[k]];
            bc[k].n = c.n;
            bc[k].nops = c.n_ops;
            bc[k].op0 = bo.size();
            bc[k].grad0 = c.grad_offset;
            nmax = std::max(nmax, c.n);
            for (int o = 0; o < c.n_ops; o++) {
                BOp b;
                int rc = lower_bop(ops[c.op_offset + o], &b);
                if (rc != SPZ_OK) {
                    release();
                    return rc;
                }
                bo.push_back(b);
            }
        }
        const uint64_t slot = 1ull << nmax;
        int occ = 0;
        SPZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_batch_adjoint, kW, 0));
        uint64_t grid = (uint64_t)rt().cu_count * std::max(occ, 1);
        grid = std::min<uint64_t>(grid, small.size());
        // scratch: 4 arrays of 2^nmax doubles per workgroup
        const uint64_t scratch_bytes = grid * 4 * slot * sizeof(double);
        for (uint64_t i : small) gmax = std::max<uint64_t>(gmax, circuits[i].grad_offset + circuits[i].n_ops);
        if (hipMalloc(&d_circ, bc.size() * sizeof(BCirc)) != hipSuccess ||
            hipMalloc(&d_ops, std::max<size_t>(bo.size(), 1) * sizeof(BOp)) != hipSuccess ||
            hipMalloc(&d_scr, scratch_bytes) != hipSuccess ||
            hipMalloc(&d_cnt, sizeof(unsigned int)) != hipSuccess ||
            hipMalloc(&d_e, small.size() * sizeof(double)) != hipSuccess ||
            hipMalloc(&d_g, std::max<uint64_t>(gmax, 1) * sizeof(double)) != hipSuccess ||
            hipMalloc(&d_t, small.size() * 2 * sizeof(unsigned long long)) != hipSuccess) {
            (void)hipGetLastError();
            cleanup();
            release();
            return fail(SPZ_ERR_OOM, "batched adjoint scratch allocation failed");
        }
        // the gradient slots of parameter-free ops stay zero
        std::vector<double> zeros(std::max<uint64_t>(gmax, 1), 0.0);
        if (hipMemcpy(d_circ, bc.data(), bc.size() * sizeof(BCirc), hipMemcpyHostToDevice) != hipSuccess ||
            (bo.size() && hipMemcpy(d_ops, bo.data(), bo.size() * sizeof(BOp), hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(d_g, zeros.data(), zeros.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            cleanup();
            release();
            return fail(SPZ_ERR_HIP, "batched adjoint upload failed");
        }
        bool ok = true;
        if (use_graph) {
            ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, (unsigned int *)d_cnt);
                hipLaunchKernelGGL(k_batch_adjoint, dim3((unsigned)grid), dim3(kW), 0, st,
                                   (const BCirc *)d_circ, (uint64_t)small.size(),
                                   (const BOp *)d_ops, hx, (double *)d_scr, slot,
                                   (unsigned int *)d_cnt, (double *)d_e, (double *)d_g,
                                   (unsigned long long *)d_t);
                ok = hipStreamEndCapture(st, &graph) == hipSuccess;
            }
            if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
            if (ok) ok = hipGraphLaunch(exec, st) == hipSuccess;
        } else {
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, (unsigned int *)d_cnt);
            hipLaunchKernelGGL(k_batch_adjoint, dim3((unsigned)grid), dim3(kW), 0, st,
                               (const BCirc *)d_circ, (uint64_t)small.size(), (const BOp *)d_ops,
                               hx, (double *)d_scr, slot, (unsigned int *)d_cnt, (double *)d_e,
                               (double *)d_g, (unsigned long long *)d_t);
            ok = hipGetLastError() == hipSuccess;
        }
        if (!ok) {
            (void)hipGetLastError();
            (void)hipStreamSynchronize(st);
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            cleanup();
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
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
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
    cleanup();
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
