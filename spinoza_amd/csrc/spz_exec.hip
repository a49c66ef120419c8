// spz_exec.hip -- QuantumCircuit::execute over an op list (circuit.rs:552-600).
#include <string.h>

#include "spz_internal.h"

namespace spz {

namespace {

int exec_one_plain(spz_state_s *s, const spz_op &o, spz_tracker *trk, uint8_t *outcome)
{
    if (outcome) *outcome = 255;
    switch (o.kind) {
    case SPZ_UNITARY: {  // transform_u / c_transform_u (circuit.rs:555-558)
        int control = -1;
        if (o.zeros) return fail(SPZ_ERR_GATE, "dense block with a 0-control");
        if (o.ones) {
            if (__builtin_popcountll(o.ones) != 1)
                return fail(SPZ_ERR_GATE, "dense block with more than one control");
            control = __builtin_ctzll(o.ones);
        }
        return spz_transform_u(s, o.m, o.u_re, o.u_im, o.target, control);
    }
    case SPZ_M: {  // circuit.rs:559-566: measure with reset unless already measured
        if (o.ones || o.zeros) return fail(SPZ_ERR_GATE, "controlled measurement");
        if (o.target < 0 || o.target >= s->n) return fail(SPZ_ERR_QUBIT, "target out of range");
        if (trk && ((trk->measured >> o.target) & 1)) return SPZ_OK;
        uint8_t v = 0;
        SPZ_TRY(measure(s, o.target, 1, -1, o.u, &v));
        if (trk) {
            trk->measured |= 1ull << o.target;
            trk->values &= ~(1ull << o.target);
            trk->values |= (uint64_t)v << o.target;
        }
        if (outcome) *outcome = v;
        return SPZ_OK;
    }
    case SPZ_BITFLIP:  // gates.rs:1365-1374
        if (o.ones || o.zeros) return fail(SPZ_ERR_GATE, "controlled bit-flip noise");
        if (o.target < 0 || o.target >= s->n) return fail(SPZ_ERR_QUBIT, "target out of range");
        if (o.u <= o.theta) return apply_gate(s, SPZ_X, 0, 0, 0, o.target, 0, 0, 0);
        return SPZ_OK;
    default:
        break;
    }
    if (o.kind < SPZ_H || o.kind > SPZ_SWAP) return fail(SPZ_ERR_GATE, "unknown gate kind");
    // Classical control once the single control was measured (circuit.rs:570-578).
    if ((o.flags & SPZ_OPF_CLASSICAL) && o.zeros == 0 && __builtin_popcountll(o.ones) == 1 && trk) {
        int c = __builtin_ctzll(o.ones);
        if ((trk->measured >> c) & 1) {
            if ((trk->values >> c) & 1)
                return apply_gate(s, o.kind, o.theta, o.phi, o.lambda, o.target, o.t1, 0, 0);
            return SPZ_OK;
        }
    }
    return apply_gate(s, o.kind, o.theta, o.phi, o.lambda, o.target, o.t1, o.ones, o.zeros);
}

}  // namespace

int execute_plain(spz_state_s *s, const spz_op *ops, uint64_t n_ops, spz_tracker *trk,
                  uint8_t *outcomes)
{
    for (uint64_t i = 0; i < n_ops; i++) {
        int rc = exec_one_plain(s, ops[i], trk, outcomes ? &outcomes[i] : nullptr);
        if (rc != SPZ_OK) {
            set_error("op " + std::to_string(i) + ": " + spz_last_error());
            return rc;
        }
    }
    return SPZ_OK;
}

int execute_ops(spz_state_s *s, const spz_op *ops, uint64_t n_ops, spz_tracker *trk,
                const spz_exec_opts *opts, uint8_t *outcomes)
{
    (void)opts;
    return execute_plain(s, ops, n_ops, trk, outcomes);
}

}  // namespace spz

extern "C" int spz_adjoint_grad(spz_state s, const spz_op *ops, uint64_t n_ops, double hx,
                                int ckpt_every, double *grad, double *energy, double *drift)
{
    (void)s; (void)ops; (void)n_ops; (void)hx; (void)ckpt_every; (void)grad; (void)energy;
    (void)drift;
    return spz::fail(SPZ_ERR_GATE, "adjoint gradient not built yet");
}
