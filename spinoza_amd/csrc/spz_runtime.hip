// spz_runtime.hip -- host runtime and the C ABI entry points.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "spz_internal.h"

namespace spz {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

static Runtime g_rt;
static std::mutex g_init_mu;

Runtime &rt() { return g_rt; }

std::recursive_mutex &api_mutex()
{
    static std::recursive_mutex *m = new std::recursive_mutex();  // lives until the process ends
    return *m;
}

int ensure_init()
{
    if (g_rt.ready) return SPZ_OK;
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_rt.ready) return SPZ_OK;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(SPZ_ERR_NODEVICE, "no HIP device visible");
    SPZ_HIP(hipGetDevice(&g_rt.device));
    hipDeviceProp_t prop;
    SPZ_HIP(hipGetDeviceProperties(&prop, g_rt.device));
    g_rt.cu_count = prop.multiProcessorCount;
    // A default-priority stream; blocking w.r.t. the null stream.
    SPZ_HIP(hipStreamCreate(&g_rt.stream));
    SPZ_HIP(hipMalloc(&g_rt.d_result, 4 * sizeof(double)));
    SPZ_HIP(hipHostMalloc(&g_rt.h_result, 4 * sizeof(double), hipHostMallocDefault));
    SPZ_TRY(static_scratch_init());
    g_rt.ready = true;
    return SPZ_OK;
}

int ensure_partials(size_t pairs)
{
    if (pairs <= g_rt.partial_cap) return SPZ_OK;
    if (g_rt.d_partials) {
        SPZ_HIP(hipStreamSynchronize(g_rt.stream));
        SPZ_HIP(hipFree(g_rt.d_partials));
        g_rt.d_partials = nullptr;
    }
    size_t cap = std::max<size_t>(pairs, 1024);
    if (hipMalloc(&g_rt.d_partials, cap * 2 * sizeof(double)) != hipSuccess)
        return fail(SPZ_ERR_OOM, "reduction scratch allocation failed");
    g_rt.partial_cap = cap;
    return SPZ_OK;
}

// ---------------------------------------------------------------------------
// Launch timing
// ---------------------------------------------------------------------------
static hipEvent_t take_event()
{
    if (!g_rt.event_pool.empty()) {
        hipEvent_t e = g_rt.event_pool.back();
        g_rt.event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

static void resolve_pending()
{
    if (g_rt.pending.empty()) return;
    hipStreamSynchronize(g_rt.stream);
    for (auto &p : g_rt.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            g_rt.launches[p.cls] += 1;
            g_rt.total_ms[p.cls] += ms;
            g_rt.alg_bytes[p.cls] += p.bytes;
        }
        g_rt.event_pool.push_back(p.a);
        g_rt.event_pool.push_back(p.b);
    }
    g_rt.pending.clear();
}

LaunchTimer::LaunchTimer(int cls_, double bytes_) : cls(cls_), bytes(bytes_)
{
    if (!g_rt.profiling) return;
    a = take_event();
    b = take_event();
    if (a && b) hipEventRecord(a, g_rt.stream);
}

LaunchTimer::~LaunchTimer()
{
    if (!a || !b) return;
    hipEventRecord(b, g_rt.stream);
    g_rt.pending.push_back({cls, a, b, bytes});
    if (g_rt.pending.size() > 16384) resolve_pending();
}

// ---------------------------------------------------------------------------
// Gate coefficients: restated from the reference's host-side set-up code
// ---------------------------------------------------------------------------
bool is_diagonal(int ak) { return ak == AK_NEG || ak == AK_PHASE || ak == AK_RZ; }

// Sine and cosine of one angle from glibc's sincos.  The reference computes
// every gate's pair with f64::sin_cos, or sin and cos of the same angle side
// by side; rustc emits LLVM's sin/cos intrinsics, which LLVM turns into one
// sincos call on Linux.  sincos and separate sin/cos calls differ by one ulp
// for some angles, so the coefficients come from sincos, as in the oracle.
static inline void sin_cos(double x, double *s, double *c) { ::sincos(x, s, c); }

int make_coef(int kind, double th, double ph, double la, Coef *k)
{
    memset(k, 0, sizeof *k);
    switch (kind) {
    case SPZ_H:  // SQRT_ONE_HALF = FRAC_1_SQRT_2 (math.rs:5)
        k->ak = AK_H;
        k->c[0] = 0.70710678118654752440;
        return SPZ_OK;
    case SPZ_X: k->ak = AK_X; return SPZ_OK;
    case SPZ_Y: k->ak = AK_Y; return SPZ_OK;
    case SPZ_Z: k->ak = AK_NEG; return SPZ_OK;
    case SPZ_P: {  // p_apply: let (sin, cos) = sin_cos(angle) (gates.rs:865)
        double s, c;
        sin_cos(th, &s, &c);
        k->ak = AK_PHASE;
        k->c[0] = c;
        k->c[1] = s;
        return SPZ_OK;
    }
    case SPZ_RX: {  // rx_apply: theta = angle*0.5; cos, -sin (gates.rs:749-751)
        double s, c;
        sin_cos(th * 0.5, &s, &c);
        k->ak = AK_RX;
        k->c[0] = c;
        k->c[1] = -s;
        return SPZ_OK;
    }
    case SPZ_RY: {  // ry_apply: (sin, cos) = sin_cos(angle*0.5) (gates.rs:1125-1126)
        double s, c;
        sin_cos(th * 0.5, &s, &c);
        k->ak = AK_RY;
        k->c[0] = s;
        k->c[1] = c;
        return SPZ_OK;
    }
    case SPZ_RZ: {  // rz_apply: d0 = (c, -s), d1 = (c, s) (gates.rs:969-975)
        double s, c;
        sin_cos(th * 0.5, &s, &c);
        k->ak = AK_RZ;
        k->c[0] = c; k->c[1] = -s; k->c[2] = c; k->c[3] = s;
        return SPZ_OK;
    }
    case SPZ_U: {  // u_apply (gates.rs:1286-1303)
        double st, ct, sl, cl, spl, cpl, sp, cp;
        sin_cos(th * 0.5, &st, &ct);
        sin_cos(la, &sl, &cl);
        sin_cos(ph + la, &spl, &cpl);
        sin_cos(ph, &sp, &cp);
        k->ak = AK_U;
        k->c[0] = ct; k->c[1] = 0.0;
        k->c[2] = -cl * st; k->c[3] = -sl * st;
        k->c[4] = cp * st; k->c[5] = sp * st;
        k->c[6] = cpl * ct; k->c[7] = spl * ct;
        return SPZ_OK;
    }
    default:
        return fail(SPZ_ERR_GATE, "gate kind has no 2x2 arithmetic");
    }
}

int check_state(const spz_state_s *s)
{
    if (!s || !s->re || !s->im) return fail(SPZ_ERR_ARG, "null state handle");
    return SPZ_OK;
}

static int check_qubits(const spz_state_s *s, int target, uint64_t ones, uint64_t zeros)
{
    if (target < 0 || target >= s->n) return fail(SPZ_ERR_QUBIT, "target qubit out of range");
    uint64_t tb = 1ull << target;
    if ((ones | zeros) & tb) return fail(SPZ_ERR_QUBIT, "target is also a control");
    if (ones & zeros) return fail(SPZ_ERR_QUBIT, "qubit both a 1-control and a 0-control");
    if (s->n < 64 && ((ones | zeros) >> s->n)) return fail(SPZ_ERR_QUBIT, "control out of range");
    return SPZ_OK;
}

int apply_gate(spz_state_s *s, int kind, double th, double ph, double la, int target, int t1,
               uint64_t ones, uint64_t zeros)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (kind == SPZ_SWAP) {
        if (target < 0 || target >= s->n || t1 < 0 || t1 >= s->n)
            return fail(SPZ_ERR_QUBIT, "SWAP qubit out of range");
        if (target == t1) return SPZ_OK;
        uint64_t sw = (1ull << target) | (1ull << t1);
        if ((ones | zeros) & sw) return fail(SPZ_ERR_QUBIT, "SWAP qubit is also a control");
        if (ones & zeros) return fail(SPZ_ERR_QUBIT, "qubit both a 1-control and a 0-control");
        return launch_swap(s, target, t1, ones, zeros);
    }
    SPZ_TRY(check_qubits(s, target, ones, zeros));
    Coef k;
    SPZ_TRY(make_coef(kind, th, ph, la, &k));
    if (is_diagonal(k.ak)) return launch_diag(s, k, target, ones, zeros);
    return launch_pair(s, k, target, ones, zeros);
}

int measure(spz_state_s *s, int target, int reset, int forced, double u, uint8_t *out)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (target < 0 || target >= s->n) return fail(SPZ_ERR_QUBIT, "target qubit out of range");
    if (forced > 1 || forced < -1) return fail(SPZ_ERR_ARG, "forced outcome must be -1, 0 or 1");
    double prob0;
    SPZ_TRY(reduce(s, nullptr, TERM_PROB0, target, 0, 0.0, &prob0));
    int val;
    if (forced >= 0) val = forced;
    else val = (u < 1.0 - prob0) ? 1 : 0;
    double p = val == 0 ? prob0 : 1.0 - prob0;
    if (!(p > 0.0)) return fail(SPZ_ERR_PROB, "measured outcome has probability 0");
    double scale = 1.0 / sqrt(p);
    SPZ_TRY(launch_collapse(s, target, val, reset, scale));
    if (out) *out = (uint8_t)val;
    return SPZ_OK;
}

}  // namespace spz

using namespace spz;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int spz_init(void) { return ensure_init(); }

const char *spz_last_error(void) { return g_last_error.c_str(); }

const char *spz_version(void) { return "spinoza_amd 0.1.0 (gfx950)"; }

int spz_device_info(char *name, size_t name_len, int *cu_count, uint64_t *hbm_bytes)
{
    SPZ_API_LOCK;
    SPZ_TRY(ensure_init());
    hipDeviceProp_t prop;
    SPZ_HIP(hipGetDeviceProperties(&prop, g_rt.device));
    if (name && name_len) {
        strncpy(name, prop.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (uint64_t)prop.totalGlobalMem;
    return SPZ_OK;
}

void *spz_stream(void)
{
    if (ensure_init() != SPZ_OK) return nullptr;
    return (void *)g_rt.stream;
}

int spz_sync(void)
{
    SPZ_API_LOCK;
    SPZ_TRY(ensure_init());
    SPZ_HIP(hipStreamSynchronize(g_rt.stream));
    return SPZ_OK;
}

int spz_state_create(int n, int precision, spz_state *out)
{
    if (!out) return fail(SPZ_ERR_ARG, "null output handle");
    *out = nullptr;
    if (n < 1 || n > 40) return fail(SPZ_ERR_QUBIT, "qubit count must be in [1, 40]");
    if (precision != SPZ_F64 && precision != SPZ_F32) return fail(SPZ_ERR_ARG, "bad precision");
    SPZ_API_LOCK;
    SPZ_TRY(ensure_init());
    auto *s = new spz_state_s();
    s->n = n;
    s->prec = precision;
    s->N = 1ull << n;
    size_t bytes = (size_t)s->N * (precision == SPZ_F64 ? 8 : 4);
    if (hipMalloc(&s->re, bytes) != hipSuccess) {
        hipGetLastError();
        delete s;
        return fail(SPZ_ERR_OOM, "state allocation failed (reals)");
    }
    if (hipMalloc(&s->im, bytes) != hipSuccess) {
        hipGetLastError();
        hipFree(s->re);
        delete s;
        return fail(SPZ_ERR_OOM, "state allocation failed (imags)");
    }
    int rc = launch_fill_basis(s, 0);
    if (rc != SPZ_OK) {
        hipFree(s->re);
        hipFree(s->im);
        delete s;
        return rc;
    }
    *out = s;
    return SPZ_OK;
}

int spz_state_destroy(spz_state s)
{
    if (!s) return SPZ_OK;
    SPZ_API_LOCK;
    if (g_rt.ready) hipStreamSynchronize(g_rt.stream);
    if (s->re) hipFree(s->re);
    if (s->im) hipFree(s->im);
    delete s;
    return SPZ_OK;
}

int spz_state_qubits(spz_state s, int *n)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!n) return fail(SPZ_ERR_ARG, "null output");
    *n = s->n;
    return SPZ_OK;
}

int spz_state_precision(spz_state s, int *p)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!p) return fail(SPZ_ERR_ARG, "null output");
    *p = s->prec;
    return SPZ_OK;
}

int spz_state_set_basis(spz_state s, uint64_t index)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (index >= s->N) return fail(SPZ_ERR_ARG, "basis index out of range");
    return launch_fill_basis(s, index);
}

static int copy_host(spz_state s, void *dst_dev, const double *src, uint64_t offset,
                     uint64_t count)
{
    if (s->prec == SPZ_F64) {
        SPZ_HIP(hipMemcpyAsync((double *)dst_dev + offset, src, count * sizeof(double),
                               hipMemcpyHostToDevice, g_rt.stream));
        SPZ_HIP(hipStreamSynchronize(g_rt.stream));
    } else {
        std::vector<float> tmp(count);
        for (uint64_t i = 0; i < count; i++) tmp[i] = (float)src[i];
        SPZ_HIP(hipMemcpyAsync((float *)dst_dev + offset, tmp.data(), count * sizeof(float),
                               hipMemcpyHostToDevice, g_rt.stream));
        SPZ_HIP(hipStreamSynchronize(g_rt.stream));
    }
    return SPZ_OK;
}

static int copy_dev(spz_state s, double *dst, const void *src_dev, uint64_t offset,
                    uint64_t count)
{
    if (s->prec == SPZ_F64) {
        SPZ_HIP(hipMemcpyAsync(dst, (const double *)src_dev + offset, count * sizeof(double),
                               hipMemcpyDeviceToHost, g_rt.stream));
        SPZ_HIP(hipStreamSynchronize(g_rt.stream));
    } else {
        std::vector<float> tmp(count);
        SPZ_HIP(hipMemcpyAsync(tmp.data(), (const float *)src_dev + offset, count * sizeof(float),
                               hipMemcpyDeviceToHost, g_rt.stream));
        SPZ_HIP(hipStreamSynchronize(g_rt.stream));
        for (uint64_t i = 0; i < count; i++) dst[i] = (double)tmp[i];
    }
    return SPZ_OK;
}

int spz_state_upload(spz_state s, const double *re, const double *im, uint64_t offset,
                     uint64_t count)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (offset > s->N || count > s->N - offset) return fail(SPZ_ERR_ARG, "range out of bounds");
    if (count == 0) return SPZ_OK;
    if (re) SPZ_TRY(copy_host(s, s->re, re, offset, count));
    if (im) SPZ_TRY(copy_host(s, s->im, im, offset, count));
    return SPZ_OK;
}

int spz_state_download(spz_state s, double *re, double *im, uint64_t offset, uint64_t count)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (offset > s->N || count > s->N - offset) return fail(SPZ_ERR_ARG, "range out of bounds");
    if (count == 0) return SPZ_OK;
    if (re) SPZ_TRY(copy_dev(s, re, s->re, offset, count));
    if (im) SPZ_TRY(copy_dev(s, im, s->im, offset, count));
    return SPZ_OK;
}

int spz_state_copy(spz_state dst, spz_state src)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(dst));
    SPZ_TRY(check_state(src));
    if (dst->n != src->n || dst->prec != src->prec)
        return fail(SPZ_ERR_ARG, "state copy between different shapes");
    if (dst == src) return SPZ_OK;
    size_t bytes = (size_t)src->N * (src->prec == SPZ_F64 ? 8 : 4);
    LaunchTimer tm(SPZ_K_OTHER, 4.0 * bytes);
    SPZ_HIP(hipMemcpyAsync(dst->re, src->re, bytes, hipMemcpyDeviceToDevice, g_rt.stream));
    SPZ_HIP(hipMemcpyAsync(dst->im, src->im, bytes, hipMemcpyDeviceToDevice, g_rt.stream));
    return SPZ_OK;
}

int spz_state_device_ptrs(spz_state s, void **re, void **im)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (re) *re = s->re;
    if (im) *im = s->im;
    return SPZ_OK;
}

int spz_apply(spz_state s, const spz_gate *g, int target, double eps)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!g) return fail(SPZ_ERR_ARG, "null gate");
    if (g->kind == SPZ_BITFLIP) {
        if (target < 0 || target >= s->n) return fail(SPZ_ERR_QUBIT, "target qubit out of range");
        if (eps <= g->theta) return apply_gate(s, SPZ_X, 0, 0, 0, target, 0, 0, 0);
        return SPZ_OK;
    }
    if (g->kind == SPZ_M || g->kind == SPZ_UNITARY)
        return fail(SPZ_ERR_GATE, "apply: M and Unitary are not single-target gates");
    return apply_gate(s, g->kind, g->theta, g->phi, g->lambda, target, g->t1, 0, 0);
}

int spz_c_apply(spz_state s, const spz_gate *g, int control, int target)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!g) return fail(SPZ_ERR_ARG, "null gate");
    if (control < 0 || control >= s->n) return fail(SPZ_ERR_QUBIT, "control out of range");
    if (g->kind > SPZ_U) return fail(SPZ_ERR_GATE, "c_apply: unsupported gate");
    return apply_gate(s, g->kind, g->theta, g->phi, g->lambda, target, g->t1, 1ull << control, 0);
}

int spz_cc_apply(spz_state s, const spz_gate *g, int c0, int c1, int target)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!g) return fail(SPZ_ERR_ARG, "null gate");
    if (c0 < 0 || c0 >= s->n || c1 < 0 || c1 >= s->n || c0 == c1)
        return fail(SPZ_ERR_QUBIT, "controls out of range or equal");
    if (g->kind > SPZ_U) return fail(SPZ_ERR_GATE, "cc_apply: unsupported gate");
    return apply_gate(s, g->kind, g->theta, g->phi, g->lambda, target, g->t1,
                      (1ull << c0) | (1ull << c1), 0);
}

int spz_mc_apply(spz_state s, const spz_gate *g, uint64_t ones, uint64_t zeros, int target)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!g) return fail(SPZ_ERR_ARG, "null gate");
    if (g->kind > SPZ_U) return fail(SPZ_ERR_GATE, "mc_apply: unsupported gate");
    return apply_gate(s, g->kind, g->theta, g->phi, g->lambda, target, g->t1, ones, zeros);
}

int spz_transform_u(spz_state s, int m, const double *u_re, const double *u_im, int t,
                    int control)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!u_re || !u_im) return fail(SPZ_ERR_ARG, "null matrix");
    if (m < 1 || m > 5) return fail(SPZ_ERR_ARG, "dense block width must be 1..5 qubits");
    if (t < 0 || t + m > s->n) return fail(SPZ_ERR_QUBIT, "block out of range");
    if (control >= 0 && (control >= s->n || (control >= t && control < t + m)))
        return fail(SPZ_ERR_QUBIT, "control must be a qubit outside the block");
    if (control < -1) return fail(SPZ_ERR_QUBIT, "bad control");
    size_t dd = (size_t)1 << (2 * m);
    double *d = nullptr;
    SPZ_HIP(hipMalloc(&d, 2 * dd * sizeof(double)));
    hipError_t e1 = hipMemcpyAsync(d, u_re, dd * sizeof(double), hipMemcpyHostToDevice, g_rt.stream);
    hipError_t e2 = hipMemcpyAsync(d + dd, u_im, dd * sizeof(double), hipMemcpyHostToDevice,
                                   g_rt.stream);
    int rc = (e1 == hipSuccess && e2 == hipSuccess) ? launch_dense(s, m, d, d + dd, t, control)
                                                    : fail(SPZ_ERR_HIP, "matrix upload failed");
    hipStreamSynchronize(g_rt.stream);
    hipFree(d);
    return rc;
}

int spz_execute(spz_state s, const spz_op *ops, uint64_t n_ops, spz_tracker *tracker,
                const spz_exec_opts *opts, uint8_t *outcomes)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (n_ops && !ops) return fail(SPZ_ERR_ARG, "null op list");
    return execute_ops(s, ops, n_ops, tracker, opts, outcomes);
}

int spz_norm(spz_state s, double *out)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!out) return fail(SPZ_ERR_ARG, "null output");
    return reduce(s, nullptr, TERM_NORM, 0, 0, 0.0, out);
}

int spz_prob0(spz_state s, int target, double *out)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!out) return fail(SPZ_ERR_ARG, "null output");
    if (target < 0 || target >= s->n) return fail(SPZ_ERR_QUBIT, "target qubit out of range");
    return reduce(s, nullptr, TERM_PROB0, target, 0, 0.0, out);
}

int spz_measure_qubit(spz_state s, int target, int reset, int forced, double u, uint8_t *out)
{
    return measure(s, target, reset, forced, u, out);
}

int spz_qubit_expectation_value(spz_state s, int target, double *out)
{
    if (!out) return fail(SPZ_ERR_ARG, "null output");
    double p0;
    SPZ_TRY(spz_prob0(s, target, &p0));
    *out = 2.0 * p0 - 1.0;  // core.rs:217-218
    return SPZ_OK;
}

int spz_xyz_expectation_value(spz_state s, char observable, const int32_t *targets,
                              uint64_t n_targets, double *out)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    int term;
    if (observable == 'x') term = TERM_X;
    else if (observable == 'y') term = TERM_Y;
    else if (observable == 'z') term = TERM_Z;
    else return fail(SPZ_ERR_ARG, std::string("observable ") + observable + " not supported");
    if (n_targets && (!targets || !out)) return fail(SPZ_ERR_ARG, "null targets/output");
    for (uint64_t i = 0; i < n_targets; i++) {
        int t = targets[i];
        if (t < 0 || t >= s->n) return fail(SPZ_ERR_QUBIT, "target qubit out of range");
        SPZ_TRY(reduce(s, nullptr, term, t, 0, 0.0, &out[i]));
    }
    return SPZ_OK;
}

int spz_expect_zz(spz_state s, int q0, int q1, double *out)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!out) return fail(SPZ_ERR_ARG, "null output");
    if (q0 < 0 || q0 >= s->n || q1 < 0 || q1 >= s->n) return fail(SPZ_ERR_QUBIT, "qubit out of range");
    return reduce(s, nullptr, TERM_ZZ, q0, q1, 0.0, out);
}

int spz_ising_energy(spz_state s, double hx, double *out)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!out) return fail(SPZ_ERR_ARG, "null output");
    return reduce(s, nullptr, TERM_ISING, 0, 0, hx, out);
}

int spz_ising_energy_noalloc(spz_state s, double hx, double *out)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (!out) return fail(SPZ_ERR_ARG, "null output");
    SPZ_TRY(ensure_init());
    return ising_energy_noalloc(s, hx, out);
}

int spz_inner(spz_state a, spz_state b, double *re, double *im)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(a));
    SPZ_TRY(check_state(b));
    if (a->n != b->n) return fail(SPZ_ERR_ARG, "inner product of states of different sizes");
    if (re) SPZ_TRY(reduce(a, b, TERM_DOT_RE, 0, 0, 0.0, re));
    if (im) SPZ_TRY(reduce(a, b, TERM_DOT_IM, 0, 0, 0.0, im));
    return SPZ_OK;
}

int spz_sample(spz_state s, const double *uniforms, uint64_t k, uint64_t *out)
{
    SPZ_API_LOCK;
    SPZ_TRY(check_state(s));
    if (k == 0) return SPZ_OK;
    if (!uniforms || !out) return fail(SPZ_ERR_ARG, "null uniforms/output");
    std::vector<double> W;
    SPZ_TRY(norm_block_partials(s, W));
    uint64_t nb = W.size();
    std::vector<double> P(nb + 1);
    P[0] = 0.0;
    int64_t lastnz = -1;
    for (uint64_t b = 0; b < nb; b++) {
        P[b + 1] = P[b] + W[b];
        if (W[b] > 0.0) lastnz = (int64_t)b;
    }
    if (lastnz < 0) return fail(SPZ_ERR_PROB, "sampling a zero state");
    double *dP = nullptr, *dU = nullptr;
    uint64_t *dO = nullptr;
    SPZ_HIP(hipMalloc(&dP, (nb + 1) * sizeof(double)));
    if (hipMalloc(&dU, k * sizeof(double)) != hipSuccess || hipMalloc(&dO, k * sizeof(uint64_t)) != hipSuccess) {
        hipFree(dP);
        if (dU) hipFree(dU);
        return fail(SPZ_ERR_OOM, "sampling scratch allocation failed");
    }
    int rc = SPZ_OK;
    if (hipMemcpyAsync(dP, P.data(), (nb + 1) * sizeof(double), hipMemcpyHostToDevice, g_rt.stream) != hipSuccess ||
        hipMemcpyAsync(dU, uniforms, k * sizeof(double), hipMemcpyHostToDevice, g_rt.stream) != hipSuccess)
        rc = fail(SPZ_ERR_HIP, "sampling upload failed");
    if (rc == SPZ_OK) rc = launch_sample(s, dP, nb, lastnz, dU, k, dO);
    if (rc == SPZ_OK &&
        hipMemcpyAsync(out, dO, k * sizeof(uint64_t), hipMemcpyDeviceToHost, g_rt.stream) != hipSuccess)
        rc = fail(SPZ_ERR_HIP, "sampling download failed");
    hipStreamSynchronize(g_rt.stream);
    hipFree(dP);
    hipFree(dU);
    hipFree(dO);
    return rc;
}

int spz_profile_enable(int on)
{
    SPZ_API_LOCK;
    SPZ_TRY(ensure_init());
    if (!on) resolve_pending();
    g_rt.profiling = on != 0;
    return SPZ_OK;
}

int spz_profile_reset(void)
{
    SPZ_API_LOCK;
    SPZ_TRY(ensure_init());
    resolve_pending();
    for (int i = 0; i < SPZ_K_COUNT; i++) {
        g_rt.launches[i] = 0;
        g_rt.total_ms[i] = 0.0;
        g_rt.alg_bytes[i] = 0.0;
    }
    return SPZ_OK;
}

int spz_profile_get(int cls, uint64_t *launches, double *total_ms, double *alg_bytes)
{
    SPZ_API_LOCK;
    SPZ_TRY(ensure_init());
    if (cls < 0 || cls >= SPZ_K_COUNT) return fail(SPZ_ERR_ARG, "bad kernel class");
    resolve_pending();
    if (launches) *launches = g_rt.launches[cls];
    if (total_ms) *total_ms = g_rt.total_ms[cls];
    if (alg_bytes) *alg_bytes = g_rt.alg_bytes[cls];
    return SPZ_OK;
}

}  // extern "C"
