/*
 * spinoza_amd.h -- C ABI of the MI355X (gfx950) state-vector engine.
 *
 * A GPU-resident replacement for the pair-amplitude state update of the
 * spinoza crate.  Every entry point below names the reference interface it
 * replaces (paths relative to the spinoza repository).  Plain C types only:
 * pointers, sizes and POD structs, no C++ or torch types.
 *
 * Conventions
 *   - Every call returns an int status: SPZ_OK (0) or one of the SPZ_ERR_*
 *     codes; spz_last_error() gives a message for the calling thread.  The
 *     reference reports errors by panicking; a host binding maps a non-zero
 *     status to its own panic/exception (see the Python package).
 *   - A spz_state owns device memory for 2^n amplitudes stored as two
 *     separate arrays (reals, imags), FP64 or FP32 -- the SoA layout of
 *     State { reals, imags, n } (core.rs:18-51).
 *   - All work on states is ordered on the library's stream (spz_stream()).
 *     One library-wide lock serialises the entry points that use shared
 *     scratch (reductions, tile programs, block fragments, adjoint and batch
 *     buffers), so calls on different handles may come from several threads.
 *     Calls on one handle still must not overlap (the reference takes
 *     &mut State): their order would be undefined.
 *   - Qubit q is bit q of the amplitude index (qubit 0 = least significant),
 *     as in the reference.
 */
#ifndef SPINOZA_AMD_H
#define SPINOZA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum spz_status {
    SPZ_OK = 0,
    SPZ_ERR_QUBIT = 1,      /* qubit index out of range / repeated        */
    SPZ_ERR_GATE = 2,       /* gate kind not supported in this position   */
    SPZ_ERR_OOM = 3,        /* device allocation failed                   */
    SPZ_ERR_HIP = 4,        /* HIP runtime error                          */
    SPZ_ERR_ARG = 5,        /* bad argument (null pointer, size, ...)     */
    SPZ_ERR_PROB = 6,       /* forced measurement of a zero-probability outcome */
    SPZ_ERR_NODEVICE = 7    /* no GPU visible                             */
};

/* ---- precision ---------------------------------------------------------- */
enum spz_precision {
    SPZ_F64 = 0,            /* Float = f64, feature "double" (math.rs:11-12) */
    SPZ_F32 = 1             /* Float = f32, feature "single" (math.rs:14-15);
                               gate arithmetic in FP64, storage in FP32,
                               reductions in FP64                            */
};

/* ---- gates: the Gate enum (gates.rs:45-74) ------------------------------ */
enum spz_gate_kind {
    SPZ_H = 0, SPZ_X = 1, SPZ_Y = 2, SPZ_Z = 3,
    SPZ_P = 4,              /* theta                                       */
    SPZ_RX = 5, SPZ_RY = 6, SPZ_RZ = 7,   /* theta                         */
    SPZ_U = 8,              /* theta, phi, lambda                          */
    SPZ_SWAP = 9,           /* SWAP(target, t1)                            */
    SPZ_BITFLIP = 10,       /* BitFlipNoise(theta = probability)           */
    SPZ_M = 11,             /* measurement (executor only)                 */
    SPZ_UNITARY = 12        /* dense 2^m x 2^m block (executor / transform) */
};

typedef struct spz_gate {
    int32_t kind;           /* spz_gate_kind                               */
    int32_t t1;             /* second qubit of SWAP                        */
    double theta, phi, lambda;
} spz_gate;

/* ---- opaque state handle ------------------------------------------------ */
typedef struct spz_state_s *spz_state;

/* ---- library ------------------------------------------------------------ */
/* Initialise the runtime on the current HIP device (called implicitly). */
int spz_init(void);
const char *spz_last_error(void);
const char *spz_version(void);
/* Device name, compute-unit count and HBM bytes of the current device. */
int spz_device_info(char *name, size_t name_len, int *cu_count, uint64_t *hbm_bytes);
/* The hipStream_t (as void*) every call is ordered on. */
void *spz_stream(void);
/* Wait for all queued work. */
int spz_sync(void);

/* ---- state: State::new / len / reals / imags (core.rs:18-51) ------------ */
/* State::new(n) (core.rs:32-42): |0...0>, 1 <= n <= 40. */
int spz_state_create(int n, int precision, spz_state *out);
int spz_state_destroy(spz_state s);
int spz_state_qubits(spz_state s, int *n);
int spz_state_precision(spz_state s, int *precision);
/* Reset to the computational basis state |index>. */
int spz_state_set_basis(spz_state s, uint64_t index);
/* Host <-> device copies of amplitudes [offset, offset+count) as FP64
 * (State.reals / State.imags are public Vecs in the reference; these are
 * the accessors a host mirror uses).  Either pointer may be NULL. */
int spz_state_upload(spz_state s, const double *re, const double *im,
                     uint64_t offset, uint64_t count);
int spz_state_download(spz_state s, double *re, double *im,
                       uint64_t offset, uint64_t count);
/* State::clone (core.rs:17 derive(Clone)): device-to-device copy; both
 * states must have the same qubit count and precision. */
int spz_state_copy(spz_state dst, spz_state src);
/* Raw device pointers of the real and imaginary arrays (for zero-copy
 * interop, e.g. torch tensors over the same memory). */
int spz_state_device_ptrs(spz_state s, void **re, void **im);

/* ---- gate application (gates.rs:206-320) -------------------------------- */
/* gates::apply(gate, &mut state, target) (gates.rs:215-232).  SWAP uses
 * (target, gate.t1).  BitFlipNoise applies X when eps <= probability (the
 * reference draws eps from thread_rng, gates.rs:1365-1374). */
int spz_apply(spz_state s, const spz_gate *g, int target, double eps);
/* gates::c_apply(gate, &mut state, control, target) (gates.rs:257-269).
 * Also accepts Z (controlled-Z), which the reference leaves as todo!(). */
int spz_c_apply(spz_state s, const spz_gate *g, int control, int target);
/* gates::cc_apply (gates.rs:272-277), any single-target gate. */
int spz_cc_apply(spz_state s, const spz_gate *g, int control0, int control1, int target);
/* gates::mc_apply (gates.rs:290-320) with explicit masks: every qubit in
 * `ones` must read 1 and every qubit in `zeros` must read 0.  (The
 * reference's mc_apply drops its `zeros` set from the mask; a binding that
 * reproduces that passes zeros = 0.) */
int spz_mc_apply(spz_state s, const spz_gate *g, uint64_t ones, uint64_t zeros, int target);
/* gates::transform_u / c_transform_u (gates.rs:1427-1487): a dense
 * 2^m x 2^m row-major complex matrix on qubits [t, t+m); control = -1 for
 * none, else a qubit outside the block that must read 1. */
int spz_transform_u(spz_state s, int m, const double *u_re, const double *u_im,
                    int t, int control);

/* ---- circuits: QuantumCircuit::execute (circuit.rs:552-600) ------------- */
enum spz_op_flags {
    SPZ_OPF_CLASSICAL = 1   /* a single control becomes classical control
                               once that qubit was measured (circuit.rs:570-578) */
};

typedef struct spz_op {
    int32_t kind;           /* spz_gate_kind                                */
    int32_t target;         /* target qubit (SWAP: first qubit; UNITARY: t) */
    double theta, phi, lambda;
    int32_t t1;             /* SWAP: second qubit                           */
    int32_t flags;          /* spz_op_flags                                 */
    uint64_t ones;          /* controls that must read 1                    */
    uint64_t zeros;         /* controls that must read 0                    */
    double u;               /* uniform draw in [0,1) for M / BITFLIP        */
    int32_t m;              /* UNITARY: block width in qubits               */
    int32_t reserved;
    const double *u_re;     /* UNITARY: 2^m x 2^m row-major, host memory    */
    const double *u_im;
} spz_op;

/* QubitTracker (circuit.rs:122-164): measured-qubit bits and their values. */
typedef struct spz_tracker {
    uint64_t measured;
    uint64_t values;
} spz_tracker;

enum spz_exec_mode {
    SPZ_EXEC_AUTO = 0,      /* fused tile sweeps where they pay off         */
    SPZ_EXEC_PLAIN = 1,     /* one kernel launch per gate, in program order */
    SPZ_EXEC_STRICT = 2,    /* fused sweeps that keep program order exactly;
                               FP64 results are bit-identical to PLAIN.  With
                               FP32 storage a sweep keeps FP64 values between
                               its gates and rounds once at the store, where
                               PLAIN rounds after every gate: the two agree
                               within FP32 rounding, not bit for bit         */
    SPZ_EXEC_DENSE5 = 3     /* runs of gates on <= 5 qubits fused into dense
                               32x32 blocks applied with MFMA (FP64 or FP32
                               to match storage); needs n >= 12             */
};

typedef struct spz_exec_opts {
    int32_t mode;           /* spz_exec_mode                                */
    int32_t tile_qubits;    /* 0 = default                                  */
    int32_t reserved[6];
} spz_exec_opts;

/* Run `ops` in order against the state.  M ops measure with reset (as the
 * reference's execute does) using op.u as the uniform draw and record the
 * outcome in *tracker (which may be NULL when no op needs it); measured
 * outcomes are also written to outcomes[i] (may be NULL; 255 when an op is
 * not a fresh measurement).  opts may be NULL (SPZ_EXEC_AUTO). */
int spz_execute(spz_state s, const spz_op *ops, uint64_t n_ops, spz_tracker *tracker,
                const spz_exec_opts *opts, uint8_t *outcomes);

/* ---- OpenQASM 2.0 loader: openqasm::load / loads (openqasm.rs:13-171) ---
 * Parse the program text src[0..len) into the ops that the reference's
 * build_circuit adds (qreg declarations; h x y z rx ry rz u on each listed
 * qubit; cp(theta) and cx with the control first; angles may use pi).
 * *n_qubits receives the total register width (registers in declaration
 * order), *n_ops the op count.  With cap = 0 (ops may be NULL) only the
 * counts are returned; otherwise ops[0..cap) must hold them all.  Errors
 * (unsupported statements, unknown registers, indices out of range) return
 * SPZ_ERR_ARG with the line in spz_last_error().  Host only: no device is
 * needed. */
int spz_qasm_parse(const char *src, uint64_t len, spz_op *ops, uint64_t cap, uint64_t *n_ops,
                   int *n_qubits);

/* ---- measurement and expectation (measurement.rs, core.rs) -------------- */
/* sum |psi|^2 with the fixed compensated FP64 tree. */
int spz_norm(spz_state s, double *out);
/* The prob0 reduction of measure_qubit (measurement.rs:16-29). */
int spz_prob0(spz_state s, int target, double *out);
/* measure_qubit(&mut state, target, reset, v) (measurement.rs:12-92).
 * forced = 0/1 is the reference's Some(v); forced = -1 draws the outcome as
 * (u < 1 - prob0) ? 1 : 0.  *out receives the outcome. */
int spz_measure_qubit(spz_state s, int target, int reset, int forced, double u, uint8_t *out);
/* qubit_expectation_value (core.rs:198-219). */
int spz_qubit_expectation_value(spz_state s, int target, double *out);
/* xyz_expectation_value(observable, &state, targets) (core.rs:222-264):
 * out[i] = Re <psi | O_{targets[i]} psi>, O in {'x','y','z'}. */
int spz_xyz_expectation_value(spz_state s, char observable, const int32_t *targets,
                              uint64_t n_targets, double *out);
/* <Z_q0 Z_q1>. */
int spz_expect_zz(spz_state s, int q0, int q1, double *out);
/* <H> for H = sum_{i<n-1} Z_i Z_{i+1} + hx * sum_i X_i (open chain). */
int spz_ising_energy(spz_state s, double hx, double *out);
/* spz_ising_energy without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device or pinned memory while they run
 * (tight serving and optimiser loops, code that forbids hipMalloc).  The
 * block partials live in a static device array of 4 MiB that the library
 * holds until the process ends, so states are limited to 2^30 amplitudes
 * (larger ones fail with SPZ_ERR_ARG).  The result is bit-identical to
 * spz_ising_energy.  The array is shared by every call in the process: the
 * library's lock serialises them, so calls from several threads do not
 * overlap on the device. */
int spz_ising_energy_noalloc(spz_state s, double hx, double *out);
/* <a|b> for two states of equal size; their storage precisions may differ
   (an FP32 state against an FP64 one: the terms are formed and summed in FP64). */
int spz_inner(spz_state a, spz_state b, double *re, double *im);
/* Replaces reservoir_sampling (core.rs:125-129): k indices drawn from
 * |psi|^2 by the inverse-CDF rule, one per uniform in [0,1). */
int spz_sample(spz_state s, const double *uniforms, uint64_t k, uint64_t *out);

/* ---- differentiation (no counterpart in the reference) ------------------ */
/* Adjoint gradient of E = <psi|H|psi>, psi = ops |psi_in>, H the Ising
 * chain above.  On return the state holds the final |psi>.  grad[i] =
 * dE/dtheta of op i for RX/RY/RZ/P ops (controlled or not) and 0 for other
 * ops.  ckpt_every > 0 stores |psi> every ckpt_every ops during the forward
 * pass and restores it in the reverse sweep; *drift (may be NULL) receives
 * the largest |psi_uncomputed - psi_checkpoint| seen. */
int spz_adjoint_grad(spz_state s, const spz_op *ops, uint64_t n_ops, double hx,
                     int ckpt_every, double *grad, double *energy, double *drift);

/* ---- executor planning statistics (no counterpart in the reference) ------
 * Host-only: how the tile executor (SPZ_EXEC_AUTO / SPZ_EXEC_STRICT) would
 * split `ops` for an n-qubit state, without touching a device.  For tuning and
 * tests of the scheduler. */
typedef struct spz_plan_info {
    uint64_t sweeps;        /* tile sweeps (one read + write of the state each) */
    uint64_t tile_gates;    /* gates applied inside sweeps                       */
    uint64_t layouts;       /* register-layout changes inside sweeps             */
    uint64_t plain_gates;   /* gates left to one-kernel-per-gate passes          */
    uint64_t lane_swaps;    /* in-wavefront layout changes (v_permlane*_swap)   */
    uint64_t frame_gates;   /* X, Y, Z absorbed into a sweep's Pauli frame (AUTO;
                               counted in tile_gates too)                        */
    uint64_t reserved[2];
} spz_plan_info;
int spz_plan_stats(int n, int precision, const spz_op *ops, uint64_t n_ops,
                   const spz_exec_opts *opts, spz_plan_info *out);

/* ---- batched differentiation (no counterpart in the reference) ---------- */
typedef struct spz_batch_circuit {
    int32_t n;              /* qubits (1..30); every circuit starts in |0...0> */
    int32_t n_ops;          /* ops of this circuit                          */
    uint64_t op_offset;     /* index of its first op in the shared op array */
    uint64_t grad_offset;   /* index of its first slot in the gradient array */
} spz_batch_circuit;

typedef struct spz_batch_opts {
    int32_t wg_max_qubits;  /* circuits with n <= this run in the persistent
                               kernel; 0 = autotuned value (or 14 before
                               spz_batch_autotune ran)                      */
    int32_t use_graph;      /* capture the persistent launch in a hipGraph  */
    int32_t group_qubits;   /* circuits in the persistent kernel with more
                               qubits run on 2^(n - group_qubits) workgroups
                               of one XCD; 0 = 12, < 0 = never (values
                               below 8 count as 8)                          */
    int32_t wg_per_cu;      /* resident persistent-kernel workgroups per CU:
                               0 = 2 (a group's members then share their CU
                               with at most one other circuit, which bounds
                               the per-circuit tail), < 0 = as many as fit  */
    int32_t reserved[4];
} spz_batch_opts;

typedef struct spz_batch_stats {
    double total_ms;        /* wall time of the whole batch                 */
    double p50_ms, p99_ms;  /* per-circuit processing time                  */
    uint64_t n_small, n_large;
    int32_t wg_max_qubits;
    int32_t n_group;        /* circuits run on a group of workgroups        */
} spz_batch_stats;

/* Energy and adjoint gradient (as spz_adjoint_grad, no checkpoints) of every
 * circuit: energies[i], grads[grad_offset + k] for op k of circuit i.  Ops
 * may be single-target gates H..U with up to 7 controls (1- or 0-controls,
 * none equal to the target, no qubit both).  Every op of every circuit is
 * checked before any work starts, by the same rule whichever path (the
 * persistent kernel or the single-circuit one) a circuit would take: one bad
 * op fails the whole call with SPZ_ERR_GATE / SPZ_ERR_QUBIT.  opts and stats
 * may be NULL. */
int spz_batch_adjoint(const spz_batch_circuit *circuits, uint64_t count, const spz_op *ops,
                      double hx, double *energies, double *grads,
                      const spz_batch_opts *opts, spz_batch_stats *stats);
/* Measure, per qubit count present in the batch, the persistent workgroup
 * kernel against the single-circuit path and keep the switch point for later
 * spz_batch_adjoint calls (also returned in *wg_max_qubits). */
int spz_batch_autotune(const spz_batch_circuit *circuits, uint64_t count, const spz_op *ops,
                       int *wg_max_qubits);

/* ---- kernel timing ------------------------------------------------------ */
enum spz_kernel_class {
    SPZ_K_PAIR = 0,         /* plain pair update (one gate per pass)       */
    SPZ_K_DIAG = 1,         /* plain diagonal update                        */
    SPZ_K_TILE = 2,         /* fused tile sweep                             */
    SPZ_K_REDUCE = 3,       /* reductions                                   */
    SPZ_K_DENSE = 4,        /* dense / MFMA block                           */
    SPZ_K_ADJOINT = 5,      /* adjoint reverse-sweep step                   */
    SPZ_K_OTHER = 6,
    SPZ_K_COUNT = 7
};
/* When enabled, every launch is bracketed by HIP events on the library
 * stream; spz_profile_get returns the launch count, the summed device time
 * in ms and the summed algorithmic HBM bytes of one kernel class. */
int spz_profile_enable(int on);
int spz_profile_reset(void);
int spz_profile_get(int kernel_class, uint64_t *launches, double *total_ms,
                    double *alg_bytes);

#ifdef __cplusplus
}
#endif
#endif /* SPINOZA_AMD_H */
