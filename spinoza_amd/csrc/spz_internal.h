// spz_internal.h -- shared declarations of the gfx950 state-vector engine.
//
// Build note: every translation unit is compiled with -ffp-contract=off, so
// each multiply and add rounds separately, as in the reference's Rust code;
// __builtin_fma marks exactly the places where the reference calls mul_add
// (gates.rs:837-838, 944-945, 1005-1016, 1256-1259).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/spinoza_amd.h"

namespace spz {

// ---------------------------------------------------------------------------
// Error handling
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define SPZ_HIP(call)                                                        \
    do {                                                                     \
        hipError_t _e = (call);                                              \
        if (_e != hipSuccess)                                                \
            return ::spz::fail(SPZ_ERR_HIP, std::string(#call) + ": " +      \
                                                hipGetErrorString(_e));      \
    } while (0)

#define SPZ_TRY(call)                                                        \
    do {                                                                     \
        int _rc = (call);                                                    \
        if (_rc != SPZ_OK) return _rc;                                       \
    } while (0)

// ---------------------------------------------------------------------------
// Runtime (one device, one stream)
// ---------------------------------------------------------------------------
struct Runtime {
    bool ready = false;
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    // reduction scratch: block partial pairs + final result
    double *d_partials = nullptr;
    size_t partial_cap = 0;      // in pairs
    double *d_result = nullptr;  // 4 doubles
    double *h_result = nullptr;  // pinned, 4 doubles
    // timing
    bool profiling = false;
    struct Pending { int cls; hipEvent_t a, b; double bytes; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    uint64_t launches[SPZ_K_COUNT] = {};
    double total_ms[SPZ_K_COUNT] = {};
    double alg_bytes[SPZ_K_COUNT] = {};
};

Runtime &rt();
int ensure_init();

// One lock serialises the entry points that use the library's shared state:
// the reduction scratch and pinned result, the tile-program ring, the dense
// block fragments, the adjoint scratch, the batched kernel's buffers and the
// timing records.  Work is ordered on the one library stream anyway, so the
// lock costs no device parallelism; it makes calls from several threads (on
// different handles) safe.  Recursive, because entry points call each other.
std::recursive_mutex &api_mutex();
#define SPZ_API_LOCK std::lock_guard<std::recursive_mutex> spz_api_lock_(::spz::api_mutex())
int ensure_partials(size_t pairs);

// Bracket a kernel launch with events when profiling is on.
struct LaunchTimer {
    int cls;
    double bytes;
    hipEvent_t a = nullptr, b = nullptr;
    LaunchTimer(int cls_, double bytes_);
    ~LaunchTimer();
};

}  // namespace spz

struct spz_state_s {
    int n = 0;
    int prec = SPZ_F64;
    uint64_t N = 0;
    void *re = nullptr;
    void *im = nullptr;
};

namespace spz {

// ---------------------------------------------------------------------------
// Gate coefficients (computed on the host with the C library's sin/cos, the
// same calls the reference's f64::sin_cos resolves to)
// ---------------------------------------------------------------------------
enum Arith : int {
    AK_H = 0,     // c0 = 1/sqrt(2)                       h_apply_strat2
    AK_X = 1,     // swap                                  x_apply_target
    AK_Y = 2,     // swap + negate                         y_proc_chunk
    AK_RX = 3,    // c0 = cos, c1 = -sin                   rx_apply_target
    AK_RY = 4,    // c0 = sin, c1 = cos                    ry_apply_strategy2
    AK_U = 5,     // g0..g3 (8 doubles)                    u_apply_target
    AK_MAT = 6,   // general 2x2 complex, row-major interleaved
    AK_NEG = 7,   // diagonal: negate (Z on |1>)           z_proc_chunk
    AK_PHASE = 8, // diagonal: multiply |1> by (c0, c1)    p_proc_chunk
    AK_RZ = 9,    // diagonal: |0>*(c0,c1), |1>*(c2,c3)    rz_apply_strategy1
    AK_USH = 10,  // fast sweeps: phase on |1>, RY, phase on |1> as shears
                  // (c0..c5: three (-tan(a/2), sin a) pairs)
};

struct Coef {
    int ak;
    int pad;
    double c[8];
};

bool is_diagonal(int ak);
// Build the arithmetic kind + coefficients of a gate (H..U).
int make_coef(int kind, double th, double ph, double la, Coef *out);

// ---------------------------------------------------------------------------
// Device pair arithmetic.  (a, b) = psi0 (re, im), (c, d) = psi1.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pair_update(int ak, const double *k, double &a, double &b,
                                            double &c, double &d)
{
    switch (ak) {
    case AK_H: {  // gates.rs:547-560
        double h = k[0];
        double a1 = h * a, b1 = h * b, c1 = h * c, d1 = h * d;
        a = a1 + c1; b = b1 + d1; c = a1 - c1; d = b1 - d1;
        break;
    }
    case AK_X: {  // gates.rs:329-334
        double t = a; a = c; c = t;
        t = b; b = d; d = t;
        break;
    }
    case AK_Y: {  // gates.rs:455-464
        double a0 = a, b0 = b;
        a = d; b = -c; c = -b0; d = a0;
        break;
    }
    case AK_RX: {  // gates.rs:721-724
        double cs = k[0], ns = k[1];
        double a0 = a, b0 = b, c0 = c, d0 = d;
        a = a0 * cs - d0 * ns;
        b = b0 * cs + c0 * ns;
        c = b0 * -ns + c0 * cs;
        d = d0 * cs + a0 * ns;
        break;
    }
    case AK_RY: {  // gates.rs:1116-1120
        double sn = k[0], cs = k[1];
        double a0 = a, b0 = b, c0 = c, d0 = d;
        a = a0 * cs - c0 * sn;
        b = b0 * cs - d0 * sn;
        c = a0 * sn + c0 * cs;
        d = b0 * sn + d0 * cs;
        break;
    }
    case AK_U: {  // gates.rs:1248-1266
        double cc = a, dd = b, m = c, n = d;
        double ga = k[0], kk = k[2], l = k[3], q = k[4], r = k[5], s = k[6], t = k[7];
        a = __builtin_fma(ga, cc, __builtin_fma(kk, m, -l * n));
        b = __builtin_fma(ga, dd, __builtin_fma(kk, n, l * m));
        c = __builtin_fma(q, cc, __builtin_fma(-r, dd, __builtin_fma(s, m, -t * n)));
        d = __builtin_fma(q, dd, __builtin_fma(r, cc, __builtin_fma(s, n, t * m)));
        break;
    }
    case AK_MAT: {  // dense 2x2, unitary_col_vec_mul order (gates.rs:1411-1415)
        double a0 = a, b0 = b, c0 = c, d0 = d;
        double r0 = 0.0, i0 = 0.0, r1 = 0.0, i1 = 0.0;
        r0 += k[0] * a0 - k[1] * b0; i0 += k[0] * b0 + k[1] * a0;
        r0 += k[2] * c0 - k[3] * d0; i0 += k[2] * d0 + k[3] * c0;
        r1 += k[4] * a0 - k[5] * b0; i1 += k[4] * b0 + k[5] * a0;
        r1 += k[6] * c0 - k[7] * d0; i1 += k[6] * d0 + k[7] * c0;
        a = r0; b = i0; c = r1; d = i1;
        break;
    }
    case AK_NEG:  // gates.rs:1054-1055
        c = -c; d = -d;
        break;
    case AK_PHASE: {  // gates.rs:835-838
        double zr = c, zi = d;
        c = __builtin_fma(zr, k[0], -zi * k[1]);
        d = __builtin_fma(zi, k[0], zr * k[1]);
        break;
    }
    case AK_RZ: {  // gates.rs:940-946
        double ar = a, ai = b, cr = c, ci = d;
        a = __builtin_fma(ar, k[0], -ai * k[1]);
        b = __builtin_fma(ar, k[1], ai * k[0]);
        c = __builtin_fma(cr, k[2], -ci * k[3]);
        d = __builtin_fma(cr, k[3], ci * k[2]);
        break;
    }
    default:
        break;
    }
}

// Diagonal update of one amplitude whose target bit is `bit`.
__device__ __forceinline__ void diag_update(int ak, const double *k, int bit, double &x, double &y)
{
    if (ak == AK_NEG) {
        if (bit) { x = -x; y = -y; }
    } else if (ak == AK_PHASE) {
        if (bit) {
            double zr = x, zi = y;
            x = __builtin_fma(zr, k[0], -zi * k[1]);
            y = __builtin_fma(zi, k[0], zr * k[1]);
        }
    } else {  // AK_RZ
        double mr = bit ? k[2] : k[0];
        double mi = bit ? k[3] : k[1];
        double cr = x, ci = y;
        x = __builtin_fma(cr, mr, -ci * mi);
        y = __builtin_fma(cr, mi, ci * mr);
    }
}

// Insert zero bits at the ascending positions pos[0..np).  The positions are
// read as whole words at a uniform index (scalar loads), never as per-lane
// byte loads, so an index costs no memory round trip per thread.
__device__ __forceinline__ uint64_t insert_zeros(uint64_t p, const int8_t *pos, int np)
{
    for (int i = 0; i < np; i++) {
        const uint32_t w = reinterpret_cast<const uint32_t *>(pos)[i >> 2];
        const int q = (int)(int8_t)(w >> (8 * (i & 3)));
        uint64_t lo = p & ((1ull << q) - 1ull);
        p = ((p >> q) << (q + 1)) | lo;
    }
    return p;
}

// Fixed positions for a launch (at most 64 distinct qubits).
struct PosList {
    int8_t pos[64];
    int np;
};

int build_poslist(uint64_t mask, PosList *pl);

// ---------------------------------------------------------------------------
// Launch helpers (implemented in the .hip files)
// ---------------------------------------------------------------------------
int launch_pair(spz_state_s *s, const Coef &k, int target, uint64_t ones, uint64_t zeros);
int launch_diag(spz_state_s *s, const Coef &k, int target, uint64_t ones, uint64_t zeros);
int launch_swap(spz_state_s *s, int q0, int q1, uint64_t ones, uint64_t zeros);
int launch_collapse(spz_state_s *s, int target, int outcome, int reset, double scale);
int launch_dense(spz_state_s *s, int m, const double *u_re, const double *u_im, int t,
                 int control);
int launch_fill_basis(spz_state_s *s, uint64_t index);

// Reductions: term kinds
enum Term : int {
    TERM_NORM = 0,
    TERM_PROB0 = 1,
    TERM_X = 2,
    TERM_Y = 3,
    TERM_Z = 4,
    TERM_ZZ = 5,
    TERM_ISING = 6,   // Re <psi | H psi>, H psi computed inline
    TERM_DOT_RE = 7,  // Re <a|b>
    TERM_DOT_IM = 8,  // Im <a|b>
};
int reduce(const spz_state_s *a, const spz_state_s *b, int term, int q0, int q1, double hx,
           double *out);
// Level-1 block partials of |psi|^2 (for sampling); returns #blocks.
int norm_block_partials(const spz_state_s *s, std::vector<double> &W);
// The Ising energy through fixed-size static scratch (no allocation); the
// scratch's device address is taken by static_scratch_init at library init.
int static_scratch_init();
int ising_energy_noalloc(const spz_state_s *s, double hx, double *out);
int launch_sample(const spz_state_s *s, const double *d_prefix, uint64_t nb, int64_t lastnz,
                  const double *d_u, uint64_t k, uint64_t *d_out);

// Generic apply of one gate (host dispatch used by the API and executor).
int apply_gate(spz_state_s *s, int kind, double th, double ph, double la, int target,
               int t1, uint64_t ones, uint64_t zeros);
int measure(spz_state_s *s, int target, int reset, int forced, double u, uint8_t *out);

void block5_fragments(const double *U, int prec, std::vector<double> &out64,
                      std::vector<float> &out32);
int launch_block5(spz_state_s *s, const int *q, const void *dfrag);

int execute_ops(spz_state_s *s, const spz_op *ops, uint64_t n_ops, spz_tracker *trk,
                const spz_exec_opts *opts, uint8_t *outcomes);

int check_state(const spz_state_s *s);

}  // namespace spz
