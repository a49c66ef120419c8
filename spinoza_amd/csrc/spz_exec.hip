// spz_exec.hip -- QuantumCircuit::execute over an op list (circuit.rs:552-600).
//
// PLAIN  : one launch per gate, in program order (the reference's loop).
// STRICT : runs of consecutive gates are packed into fused tile sweeps
//          without reordering anything, so FP64 results are bit-identical to
//          PLAIN (FP32 storage: a sweep rounds to FP32 once at its store,
//          PLAIN after every gate, so they agree within FP32 rounding).
// AUTO   : like STRICT, but a gate may be pulled ahead of earlier gates it
//          commutes with (disjoint qubits, or qubits on which both act
//          diagonally: Z-type targets and controls), which packs more gates
//          into each sweep; floating-point results then differ from PLAIN only
//          by rounding (the same products in a different order).
//
// Measurements, SWAPs and dense blocks end a run and execute on their own.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <cmath>
#include <vector>

#include "spz_internal.h"
#include "spz_tile.h"

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

// A single-target gate of a fusable run.
struct SG {
    int op_index;
    int kind;
    int target;
    uint64_t ones, zeros;
    Coef k;
    bool diag;
    uint64_t nd;  // qubits acted on non-diagonally (the target of a 2x2 gate)
    uint64_t dm;  // qubits acted on diagonally (controls, diagonal targets)
};

bool is_barrier(const spz_op &o)
{
    return o.kind == SPZ_M || o.kind == SPZ_SWAP || o.kind == SPZ_UNITARY || o.kind < SPZ_H ||
           o.kind > SPZ_BITFLIP;
}

// Validate and lower one op of a run to an SG; returns 1 if it is a no-op.
int lower(const spz_state_s *s, const spz_op &o, const spz_tracker *trk, int idx, SG *g)
{
    int kind = o.kind;
    uint64_t ones = o.ones, zeros = o.zeros;
    if (kind == SPZ_BITFLIP) {
        if (ones || zeros) return fail(SPZ_ERR_GATE, "controlled bit-flip noise");
        if (!(o.u <= o.theta)) return -1;  // no flip: nothing to do
        kind = SPZ_X;
    }
    if ((o.flags & SPZ_OPF_CLASSICAL) && zeros == 0 && __builtin_popcountll(ones) == 1 && trk) {
        int c = __builtin_ctzll(ones);
        if ((trk->measured >> c) & 1) {
            if (!((trk->values >> c) & 1)) return -1;
            ones = 0;
        }
    }
    const int t = o.target;
    if (t < 0 || t >= s->n) return fail(SPZ_ERR_QUBIT, "target qubit out of range");
    const uint64_t tb = 1ull << t;
    if ((ones | zeros) & tb) return fail(SPZ_ERR_QUBIT, "target is also a control");
    if (ones & zeros) return fail(SPZ_ERR_QUBIT, "qubit both a 1-control and a 0-control");
    if (s->n < 64 && ((ones | zeros) >> s->n)) return fail(SPZ_ERR_QUBIT, "control out of range");
    g->op_index = idx;
    g->kind = kind;
    g->target = t;
    g->ones = ones;
    g->zeros = zeros;
    SPZ_TRY(make_coef(kind, o.theta, o.phi, o.lambda, &g->k));
    g->diag = is_diagonal(g->k.ak);
    g->nd = g->diag ? 0 : tb;
    g->dm = (ones | zeros) | (g->diag ? tb : 0);
    return SPZ_OK;
}

int env_int(const char *name, int dflt)
{
    const char *e = getenv(name);
    return (e && *e) ? atoi(e) : dflt;
}

struct Sweep {
    bool tile;                 // false: a single gate on the plain path
    int gate;                  // plain: index into the run
    TileDesc desc;
    size_t op_begin, op_count; // range in the program buffer
    // fast sweeps: the frame a next tile sweep can take over (Z part and
    // phase; the X part is always applied at the store)
    bool can_carry;
    uint64_t carry_z;
    int carry_w;
};

// One-hot TileGroup bit of an arithmetic kind (0: not a tile kind).  AK_MAT
// comes only from a fast sweep's frame conjugation (lower() never makes it).
uint32_t kind_bit(int ak)
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
    case AK_MAT: return TG_MAT;
    case AK_USH: return TG_USH;
    default: return 0;
    }
}

// A change of register layout inside a sweep: an LDS transpose to any
// layout, or in-wavefront swaps of register-index bits with lane bits 4 and 5
// (v_permlane16/32_swap), which move no data through LDS.
struct Change {
    bool lds = true;
    std::vector<int> rs;                    // register slot -> local bit after it
    std::vector<int> thr;                   // thread-index bit -> local bit after it
    std::vector<std::pair<int, int>> sw;    // (register slot, lane bit), in order
};

void gate_matrix2(const Coef &k, double m[8]);

// Pauli frame of a fast sweep: the registers hold phi, and the state is
// psi = i^w X^x Z^z phi (X^x flips the qubits of x, Z^z negates amplitudes of
// odd parity on z).  Uncontrolled X, Y and Z change only the frame; H, CX and
// CZ run unchanged and carry the frame through (Clifford conjugation); every
// other gate runs conjugated by the frame, which leaves the frame as it is.
// The store applies it (see TileDesc).
struct Frame {
    uint64_t x = 0, z = 0;
    int w = 0;
};

// Coefficients of Z^b X^a G X^a Z^b: gate G seen through a frame with X^a Z^b
// on its target.  Keeps the kind where one covers the result (RX, RY, RZ, U
// under Z; a diagonal stays diagonal), else a general matrix (AK_MAT).
void conj_coef(Coef &k, bool a, bool b)
{
    if (!a && !b) return;
    switch (k.ak) {
    case AK_RX:  // [[c, -is], [-is, c]]: X keeps it, Z negates s
        if (b) k.c[1] = -k.c[1];
        return;
    case AK_RY:  // [[c, -s], [s, c]]: X and Z each negate s
        if (a != b) k.c[0] = -k.c[0];
        return;
    case AK_RZ:
        if (a) {
            std::swap(k.c[0], k.c[2]);
            std::swap(k.c[1], k.c[3]);
        }
        return;
    case AK_PHASE:  // X diag(1, p) X = diag(p, 1)
        if (a) {
            const double pr = k.c[0], pi = k.c[1];
            memset(k.c, 0, sizeof k.c);
            k.ak = AK_RZ;
            k.c[0] = pr; k.c[1] = pi; k.c[2] = 1.0;
        }
        return;
    case AK_NEG:  // X Z X = diag(-1, 1)
        if (a) {
            memset(k.c, 0, sizeof k.c);
            k.ak = AK_RZ;
            k.c[0] = -1.0; k.c[2] = 1.0;
        }
        return;
    case AK_U:  // Z negates the off-diagonal entries; X needs a complex u00
        if (!a) {
            for (int i = 2; i < 6; i++) k.c[i] = -k.c[i];
            return;
        }
        break;
    default:
        break;
    }
    double m[8];
    gate_matrix2(k, m);
    if (a) {  // X G X: rows and columns swapped
        const double r[8] = {m[6], m[7], m[4], m[5], m[2], m[3], m[0], m[1]};
        memcpy(m, r, sizeof r);
    }
    if (b)
        for (int i = 2; i < 6; i++) m[i] = -m[i];
    k.ak = AK_MAT;
    memcpy(k.c, m, sizeof m);
}

// The fast sweeps' rotation and phase forms (shears): a rotation by a, |a| <=
// pi/2, given by (cos a, sin a) with cos a >= 0, becomes (-tan(a/2), sin a).
// RX(theta) and RY(theta) rotate by a = theta/2; a phase e^(ib) (P, and each
// half of a controlled general diagonal) by b.  A larger angle a is a - pi
// and a sign: an uncontrolled RX/RY then runs as -G (sign 1: the caller adds
// a global -1), an uncontrolled P(b) as Z P(b - pi) (sign 2: the caller puts
// a Z on the target after it); a controlled one runs twice at half the angle
// (returns 2).  Other kinds are left as they are (returns 1).
void shear_of(double c, double s, double *k)
{
    k[0] = -s / (1.0 + c);
    k[1] = s;
}

// e^(ia/2), cos >= 0, for e^(ia) = (c, s), a in (-pi, pi]: the larger of
// cos(a/2) and |sin(a/2)| from (1 + c) / 2 or (1 - c) / 2, the other from
// sin a = 2 sin(a/2) cos(a/2), so that neither cancels.
void half_of(double c, double s, double *hc, double *hs)
{
    if (c >= 0.0) {
        const double h = std::sqrt(0.5 * (1.0 + c));
        *hc = h;
        *hs = s / (2.0 * h);
    } else {
        const double q = std::sqrt(0.5 * (1.0 - c));
        *hc = std::fabs(s) / (2.0 * q);
        *hs = std::signbit(s) ? -q : q;
    }
}

int shear_form(Coef &k, bool unc, int *sign)
{
    *sign = 0;
    double c, s;
    switch (k.ak) {
    case AK_RX: c = k.c[0]; s = -k.c[1]; break;
    case AK_RY: c = k.c[1]; s = k.c[0]; break;
    case AK_PHASE: c = k.c[0]; s = k.c[1]; break;
    case AK_RZ: {  // controlled general diagonal: both halves
        double c0 = k.c[0], s0 = k.c[1], c1 = k.c[2], s1 = k.c[3];
        const int reps = (c0 < 0.0 || c1 < 0.0) ? 2 : 1;
        if (reps == 2) {
            half_of(c0, s0, &c0, &s0);
            half_of(c1, s1, &c1, &s1);
        }
        memset(k.c, 0, sizeof k.c);
        shear_of(c0, s0, k.c);
        shear_of(c1, s1, k.c + 2);
        return reps;
    }
    default: return 1;
    }
    int reps = 1;
    if (c < 0.0) {
        if (unc) {
            c = -c;
            s = -s;
            *sign = k.ak == AK_PHASE ? 2 : 1;
        } else {
            half_of(c, s, &c, &s);
            reps = 2;
        }
    }
    memset(k.c, 0, sizeof k.c);
    shear_of(c, s, k.c);
    return reps;
}

// An uncontrolled 2x2 unitary M (U, or a gate the frame turned into a general
// matrix) as e^(i al) P(be) RY(2a) P(de), P(x) = diag(1, e^(ix)): |m00| = cos a,
// |m10| = sin a, a in [0, pi/2], and the phases from whichever entries are
// the larger (so a near-zero sin or cos does not amplify rounding).  For the
// shears both phases must have |angle| <= pi/2: P(de) = Z P(de - pi) and
// Z RY(2a) Z = RY(-2a) turn (be, a, de) into (be + pi, -a, de - pi), and a
// remaining P(be) becomes Z P(be - pi), the Z going on phi after the op
// (*zafter).  k gets AK_USH with (-tan(x/2), sin x) for de, a and be; the
// global phase e^(i al) is returned in (*gr, *gi).
void ush_form(Coef &k, double *gr, double *gi, bool *zafter)
{
    double m[8];
    gate_matrix2(k, m);
    const double ca = std::hypot(m[0], m[1]), sa = std::hypot(m[4], m[5]);
    double ar, ai, br, bi, dr, di;  // e^(i al), e^(i (al + be)), e^(i (al + de))
    if (ca >= sa) {
        ar = m[0] / ca;
        ai = m[1] / ca;
        const double tr = m[6] / ca, ti = m[7] / ca;  // e^(i (al + be + de))
        if (sa > 0.0) {
            br = m[4] / sa;
            bi = m[5] / sa;
        } else {
            br = ar;
            bi = ai;
        }
        const double xr = tr * br + ti * bi, xi = ti * br - tr * bi;  // e^(i de)
        dr = xr * ar - xi * ai;
        di = xr * ai + xi * ar;
    } else {
        br = m[4] / sa;
        bi = m[5] / sa;
        dr = -m[2] / sa;
        di = -m[3] / sa;
        if (ca > 0.0) {
            ar = m[0] / ca;
            ai = m[1] / ca;
        } else {
            ar = 1.0;
            ai = 0.0;
        }
    }
    // e^(i be), e^(i de): the conjugate of e^(i al) taken out
    double ebr = br * ar + bi * ai, ebi = bi * ar - br * ai;
    double edr = dr * ar + di * ai, edi = di * ar - dr * ai;
    double s = sa;
    if (edr < 0.0) {
        edr = -edr;
        edi = -edi;
        ebr = -ebr;
        ebi = -ebi;
        s = -s;
    }
    *zafter = ebr < 0.0;
    if (*zafter) {
        ebr = -ebr;
        ebi = -ebi;
    }
    memset(k.c, 0, sizeof k.c);
    k.ak = AK_USH;
    shear_of(edr, edi, k.c);
    shear_of(ca, s, k.c + 2);
    shear_of(ebr, ebi, k.c + 4);
    *gr = ar;
    *gi = ai;
}

struct Planner {
    int n, K, LOGW, RB, klo;
    int improve = 0;  // local-search passes over the tile set (AUTO only)
    bool strict;
    bool fast = false;   // AUTO arithmetic (see launch_tile): H factors carried
    bool swaps = false;  // in-wavefront layout changes when they reach the set
    bool bad = false;
    uint64_t frame_gates = 0;  // gates absorbed into Pauli frames
    uint64_t split_ops = 0;    // second halves of gates run as two half-angle ops
    std::vector<TileOp> prog;
    std::vector<Sweep> sweeps;

    // A gate that a fast sweep absorbs into its frame needs no register slot.
    bool absorbed(const SG &g) const
    {
        return fast && !g.ones && !g.zeros &&
               (g.k.ak == AK_X || g.k.ak == AK_Y || g.k.ak == AK_NEG);
    }
    // Gates that run only with their target on a register bit.
    bool needs_reg(const SG &g) const { return !g.diag && !absorbed(g); }

    // Choose the tile qubits and the gates of one sweep from `rem` (indices
    // into `run`, in program order).  Returns the chosen gate indices.
    std::vector<int> collect(const std::vector<SG> &run, const std::vector<int> &rem,
                             uint64_t &smask)
    {
        std::vector<int> order;
        smask = (klo >= 64) ? ~0ull : ((1ull << klo) - 1);
        int free_slots = K - klo;
        uint64_t bnd = 0, bd = 0;
        const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
        for (int idx : rem) {
            const SG &g = run[idx];
            bool ok = !(g.nd & (bnd | bd)) && !(g.dm & bnd);
            if (ok && g.nd && !(g.nd & smask)) {
                if (free_slots > 0) {
                    smask |= g.nd;
                    free_slots--;
                } else {
                    ok = false;
                }
            }
            if (ok) {
                order.push_back(idx);
            } else {
                if (strict) break;
                bnd |= g.nd;
                bd |= g.dm;
                if ((bnd & all) == all) break;
            }
        }
        for (int q = 0; q < n && __builtin_popcountll(smask) < K; q++) smask |= 1ull << q;
        if (!strict && improve > 0 && n > K) improve_set(run, rem, smask, order);
        return order;
    }

    // Gates of `rem` that one sweep over the fixed tile set `smask` can take
    // (program order, skipping past blocked gates); stops once every tile
    // qubit is blocked.  With `out`, also lists them.
    size_t take_fixed(const std::vector<SG> &run, const std::vector<int> &rem, uint64_t smask,
                      std::vector<int> *out) const
    {
        uint64_t bnd = 0, bd = 0;
        size_t cnt = 0;
        for (int idx : rem) {
            const SG &g = run[idx];
            if (!(g.nd & (bnd | bd)) && !(g.dm & bnd) && !(g.nd & ~smask)) {
                cnt++;
                if (out) out->push_back(idx);
            } else {
                bnd |= g.nd;
                bd |= g.dm;
                if ((bnd & smask) == smask) break;
            }
        }
        return cnt;
    }

    // Local search over the tile set: swap a free tile qubit for a qubit that
    // the next gates target, keep the first swap that lets the sweep take more
    // gates, repeat.  The low klo qubits stay (they make the tile rows contiguous).
    void improve_set(const std::vector<SG> &run, const std::vector<int> &rem, uint64_t &smask,
                     std::vector<int> &order) const
    {
        const uint64_t fixed = (klo >= 64) ? ~0ull : ((1ull << klo) - 1);
        uint64_t cand = 0;  // non-diagonal targets among the next gates
        size_t seen = 0;
        for (int idx : rem) {
            cand |= run[idx].nd;
            if (++seen >= 4 * (size_t)n * 8) break;
        }
        cand &= ~fixed;
        size_t best = order.size();
        bool moved = false;
        for (int pass = 0; pass < improve; pass++) {
            bool better = false;
            for (int a = 0; a < n && !better; a++) {
                if (!((smask >> a) & 1) || ((fixed >> a) & 1)) continue;
                for (int b = 0; b < n && !better; b++) {
                    if (((smask >> b) & 1) || !((cand >> b) & 1)) continue;
                    const uint64_t m2 = (smask & ~(1ull << a)) | (1ull << b);
                    const size_t c2 = take_fixed(run, rem, m2, nullptr);
                    if (c2 > best) {
                        best = c2;
                        smask = m2;
                        better = moved = true;
                    }
                }
            }
            if (!better) break;
        }
        if (moved) {
            order.clear();
            take_fixed(run, rem, smask, &order);
        }
    }

    // Passes a greedy plan of `rem` would make: one per sweep, one half per
    // gate left to the plain path.
    double tail_cost(const std::vector<SG> &run, std::vector<int> rem)
    {
        double cost = 0.0;
        while (!rem.empty()) {
            uint64_t sm;
            const std::vector<int> ord = collect(run, rem, sm);
            if (ord.size() < 2) {
                cost += 0.5;
                rem.erase(rem.begin());
                continue;
            }
            cost += 1.0;
            rem = minus(rem, ord);
        }
        return cost;
    }

    // rem without the gates of ord (ord: a subsequence of rem).
    static std::vector<int> minus(const std::vector<int> &rem, const std::vector<int> &ord)
    {
        std::vector<int> out;
        out.reserve(rem.size());
        size_t k = 0;
        for (int idx : rem) {
            if (k < ord.size() && ord[k] == idx) { k++; continue; }
            out.push_back(idx);
        }
        return out;
    }

    // End of a run: a greedy plan can leave one gate that no sweep takes, a
    // whole plain pass for one gate.  When about one sweep's worth of gates
    // remains after this sweep and its plan would end that way, the tile sets
    // one swap away from this sweep's are tried, and one whose remainder
    // plans in fewer passes is kept.  Only for states of 2^28 amplitudes and
    // more, where a pass takes milliseconds and this search overlaps the
    // sweeps already queued.
    void endgame(const std::vector<SG> &run, const std::vector<int> &rem, uint64_t &smask,
                 std::vector<int> &order)
    {
        if (strict || improve <= 0 || n < 28 || n <= K) return;
        if (order.size() >= rem.size() || rem.size() - order.size() > order.size()) return;
        const double base = tail_cost(run, minus(rem, order));
        if (base == std::floor(base)) return;  // no plain gate in the tail
        const uint64_t fixed = (klo >= 64) ? ~0ull : ((1ull << klo) - 1);
        uint64_t cand = 0;
        for (int idx : rem) cand |= run[idx].nd;
        cand &= ~fixed;
        double best = base;
        uint64_t bm = smask;
        std::vector<int> bo;
        for (int a = 0; a < n; a++) {
            if (!((smask >> a) & 1) || ((fixed >> a) & 1)) continue;
            for (int b = 0; b < n; b++) {
                if (((smask >> b) & 1) || !((cand >> b) & 1)) continue;
                const uint64_t m2 = (smask & ~(1ull << a)) | (1ull << b);
                std::vector<int> o2;
                take_fixed(run, rem, m2, &o2);
                if (o2.size() < 2) continue;
                const double c = tail_cost(run, minus(rem, o2));
                if (c < best) {
                    best = c;
                    bm = m2;
                    bo.swap(o2);
                }
            }
        }
        if (best < base) {
            smask = bm;
            order.swap(bo);
        }
    }

    void emit_layout(const std::vector<int> &oldRS, const std::vector<int> &RSv,
                     const std::vector<int> &thr)
    {
        TileOp op;
        memset(&op, 0, sizeof op);
        op.type = TOP_LAYOUT;
        op.grp = TG_LAYOUT;
        op.eo = 1;  // em = 0: never taken for a gate
        for (int j = 0; j < RB; j++) {
            op.lay.oldreg[j] = (int8_t)oldRS[j];
            op.lay.newreg[j] = (int8_t)RSv[j];
            op.lay.rob[j] = tile_swz(1u << oldRS[j]) << 3;
            op.lay.rnb[j] = tile_swz(1u << RSv[j]) << 3;
        }
        for (int i = 0; i < LOGW; i++) op.lay.newthr[i] = (int8_t)thr[i];
        prog.push_back(op);
    }

    // Register slot j trades its local bit with the one on lane bit tb (4 or 5).
    void emit_swap(int j, int tb, int reg_bit, int lane_bit)
    {
        TileOp op;
        memset(&op, 0, sizeof op);
        op.type = TOP_LAYOUT;
        op.grp = tb == 5 ? TG_SW32 : TG_SW16;
        op.eo = 1;  // never taken for a gate
        op.sel = 1u << j;
        op.mlo = 1u << lane_bit;  // leaves the lane part
        op.mlz = 1u << reg_bit;   // enters it
        prog.push_back(op);
    }

    // Thread layout after an LDS change to register bits rs: lane bits 4 and
    // 5 take the two best-scored free bits (the next register candidates, so
    // that later changes can be swaps); the other free bits fill the thread
    // bits in ascending order, low bits first.  (lo: bits below it may not
    // go to lanes 4, 5, used for the load layout, whose lanes 0..klo-1 must
    // hold the low bits.)
    std::vector<int> thread_layout(const std::vector<int> &rs, const std::vector<long> *score,
                                   int lo) const
    {
        std::vector<int> avail;
        for (int b = 0; b < K; b++)
            if (std::find(rs.begin(), rs.end(), b) == rs.end()) avail.push_back(b);
        std::vector<int> thr(LOGW, -1);
        if (swaps && LOGW >= 6 && score) {
            std::vector<int> cand;
            for (int b : avail)
                if (b >= lo && (*score)[b] > 0) cand.push_back(b);
            std::stable_sort(cand.begin(), cand.end(),
                             [&](int x, int y) { return (*score)[x] > (*score)[y]; });
            for (int i = 0; i < 2 && i < (int)cand.size(); i++) {
                thr[4 + i] = cand[i];
                avail.erase(std::find(avail.begin(), avail.end(), cand[i]));
            }
        }
        size_t a = 0;
        for (int i = 0; i < LOGW; i++)
            if (thr[i] < 0) thr[i] = avail[a++];
        return thr;
    }

    // Swaps that turn register set rs (slots) into one holding exactly the
    // bits of `want`, if every missing bit sits on lane bit 4 or 5.
    bool swap_path(const std::vector<int> &rs, const std::vector<int> &thr,
                   const std::vector<int> &want, std::vector<std::pair<int, int>> *sw) const
    {
        if (!swaps || LOGW < 6) return false;
        std::vector<int> in, out;
        for (int b : want)
            if (std::find(rs.begin(), rs.end(), b) == rs.end()) in.push_back(b);
        for (int j = 0; j < RB; j++)
            if (std::find(want.begin(), want.end(), rs[j]) == want.end()) out.push_back(j);
        if (in.size() != out.size() || in.size() > 2) return false;
        for (size_t i = 0; i < in.size(); i++) {
            const int tb = thr[4] == in[i] ? 4 : thr[5] == in[i] ? 5 : -1;
            if (tb < 0) return false;
            if (sw) sw->push_back({out[i], tb});
        }
        return true;
    }

    void apply_swaps(std::vector<int> &rs, std::vector<int> &thr,
                     const std::vector<std::pair<int, int>> &sw) const
    {
        for (auto [j, tb] : sw) std::swap(rs[j], thr[tb]);
    }

    // Dependency graph of a sweep's gates: a -> b (a before b) when they share
    // a qubit on which at least one acts non-diagonally.
    void deps_of(const std::vector<SG> &run, const std::vector<int> &order,
                 std::vector<std::vector<int>> &succ, std::vector<int> &indeg) const
    {
        const int m = (int)order.size();
        succ.assign(m, {});
        indeg.assign(m, 0);
        for (int a = 0; a < m; a++) {
            const SG &x = run[order[a]];
            for (int b = a + 1; b < m; b++) {
                const SG &y = run[order[b]];
                if ((x.nd & (y.nd | y.dm)) || (x.dm & y.nd)) {
                    succ[a].push_back(b);
                    indeg[b]++;
                }
            }
        }
    }

    // Gates that run, in dependency order, under register bits rs before a
    // pair gate on another bit blocks every remaining one.
    int runs_under(const std::vector<SG> &run, const std::vector<int> &order, const int *lbit,
                   const std::vector<std::vector<int>> &succ, std::vector<int> indeg,
                   std::vector<char> done, uint32_t rs) const
    {
        const int m = (int)order.size();
        int cnt = 0;
        bool progress = true;
        while (progress) {
            progress = false;
            for (int a = 0; a < m; a++) {
                if (done[a] || indeg[a]) continue;
                const SG &g = run[order[a]];
                if (needs_reg(g) && !((rs >> lbit[g.target]) & 1)) continue;
                done[a] = 1;
                cnt++;
                for (int b : succ[a]) indeg[b]--;
                progress = true;
            }
        }
        return cnt;
    }

    // The RB register bits (from `lo` up) to switch to: every RB-subset of
    // the (up to 8) best-scored bits is tried, and the one under which the
    // most gates run wins (ties: the higher score, then the earlier subset).
    // With a current layout (cur_rs, cur_thr), a set reachable by in-wavefront
    // swaps is charged less than one that needs an LDS transpose (in gates
    // that could have run: a transpose costs about four gates, a swap about
    // half of one).
    std::vector<int> best_set(const std::vector<SG> &run, const std::vector<int> &order,
                              const int *lbit, const std::vector<std::vector<int>> &succ,
                              const std::vector<int> &indeg, const std::vector<char> &done,
                              const std::vector<long> &score, int lo,
                              const std::vector<int> *cur_rs = nullptr,
                              const std::vector<int> *cur_thr = nullptr) const
    {
        int left = 0;
        for (char d : done) left += !d;
        std::vector<int> bits;
        for (int b = lo; b < K; b++) bits.push_back(b);
        std::stable_sort(bits.begin(), bits.end(),
                         [&](int x, int y) { return score[x] > score[y]; });
        const int T = std::min((int)bits.size(), std::max(RB, 8));
        uint32_t best = 0;
        double best_val = -1e30;
        long best_sc = -1;
        for (uint32_t sub = 0; sub < (1u << T); sub++) {
            if (__builtin_popcount(sub) != RB) continue;
            uint32_t rs = 0;
            long sc = 0;
            std::vector<int> want;
            for (int i = 0; i < T; i++)
                if ((sub >> i) & 1) {
                    rs |= 1u << bits[i];
                    sc += score[bits[i]];
                    want.push_back(bits[i]);
                }
            int cnt = runs_under(run, order, lbit, succ, indeg, done, rs);
            // a set that finishes the sweep and keeps the low bits on lanes
            // saves the change back before the store
            if (cnt == left && !(rs & ((1u << klo) - 1))) cnt++;
            // a set under which nothing runs makes no progress
            if (cnt == 0) continue;
            double val = cnt;
            if (cur_rs && cur_thr) {
                std::vector<std::pair<int, int>> sw;
                val -= swap_path(*cur_rs, *cur_thr, want, &sw) ? 0.5 * (double)sw.size() : 4.0;
            }
            if (val > best_val || (val == best_val && sc > best_sc)) {
                best = rs;
                best_val = val;
                best_sc = sc;
            }
        }
        std::vector<int> out;
        if (best_val == -1e30) {
            // nothing runs under any candidate (cannot happen while a ready
            // pair gate's target is among the best-scored bits): take the top
            // bits, which hold every ready target
            for (int i = 0; i < RB && i < (int)bits.size(); i++) out.push_back(bits[i]);
            std::sort(out.begin(), out.end());
            return out;
        }
        for (int b = 0; b < K; b++)
            if ((best >> b) & 1) out.push_back(b);
        return out;
    }

    // Sequence the sweep's gates: the scheduled order itself (strict), or,
    // with reordering allowed, a topological order of the dependency graph
    // that applies every gate the current register layout allows before
    // switching layouts.  Entries >= 0 are gates; -1 marks "apply the next
    // entry of `changes`".  rs_in / thr_in receive the load layout: register
    // bits never among the klo low local bits, which stay on the low lane bits
    // for contiguous loads (with reordering allowed, the register bits with
    // the most ready pair gates, and lane bits 4 and 5 holding the next
    // best-scored bits).
    void sequence(const std::vector<SG> &run, const std::vector<int> &order, const int *lbit,
                  std::vector<int> &seq, std::vector<Change> &changes, std::vector<int> &rs_in,
                  std::vector<int> &thr_in)
    {
        const int m = (int)order.size();
        std::vector<int> RS(RB);
        for (int j = 0; j < RB; j++) RS[j] = LOGW + j;
        auto in_rs = [&](int b) {
            for (int j = 0; j < RB; j++)
                if (RS[j] == b) return true;
            return false;
        };
        // pair-gate targets still to run, weighted by readiness
        auto pair_score = [&](const std::vector<int> &indeg, const std::vector<char> *done) {
            std::vector<long> score(K, 0);
            for (int a = 0; a < m; a++) {
                if (done && (*done)[a]) continue;
                const SG &g = run[order[a]];
                if (needs_reg(g)) score[lbit[g.target]] += indeg[a] == 0 ? 1000 : 1;
            }
            return score;
        };
        std::vector<std::vector<int>> succ0;
        std::vector<int> indeg0;
        deps_of(run, order, succ0, indeg0);
        std::vector<long> score0 = pair_score(indeg0, nullptr);
        if (!strict && K - klo >= RB)
            RS = best_set(run, order, lbit, succ0, indeg0, std::vector<char>(m, 0), score0, klo);
        std::vector<int> THR = thread_layout(RS, &score0, klo);
        rs_in = RS;
        thr_in = THR;
        // move to register bits `nx` (slot order as given): swaps when they
        // reach it, else an LDS transpose whose lane bits 4, 5 get the
        // best-scored free bits
        auto change_to = [&](const std::vector<int> &nx, const std::vector<long> &score) {
            Change ch;
            if (swap_path(RS, THR, nx, &ch.sw)) {
                ch.lds = false;
                apply_swaps(RS, THR, ch.sw);
            } else {
                // keep bits that stay in place in their old slots
                std::vector<int> slotted(RB, -1);
                for (int j = 0; j < RB; j++)
                    if (std::find(nx.begin(), nx.end(), RS[j]) != nx.end()) slotted[j] = RS[j];
                for (int b : nx) {
                    if (std::find(slotted.begin(), slotted.end(), b) != slotted.end()) continue;
                    for (int j = 0; j < RB; j++)
                        if (slotted[j] < 0) { slotted[j] = b; break; }
                }
                RS = slotted;
                THR = thread_layout(RS, &score, 0);
            }
            ch.rs = RS;
            ch.thr = THR;
            seq.push_back(-1);
            changes.push_back(ch);
        };
        if (strict) {
            for (int pos = 0; pos < m; pos++) {
                const SG &g = run[order[pos]];
                if (!g.diag && !in_rs(lbit[g.target])) {
                    std::vector<int> nx;
                    std::vector<long> score(K, 0);
                    for (int p2 = pos; p2 < m; p2++) {
                        const SG &h = run[order[p2]];
                        if (h.diag) continue;
                        int hb = lbit[h.target];
                        score[hb] += std::max(1, 64 - (p2 - pos));
                        if ((int)nx.size() < RB && std::find(nx.begin(), nx.end(), hb) == nx.end())
                            nx.push_back(hb);
                    }
                    for (int j = 0; j < RB && (int)nx.size() < RB; j++)
                        if (std::find(nx.begin(), nx.end(), RS[j]) == nx.end()) nx.push_back(RS[j]);
                    for (int bb = 0; bb < K && (int)nx.size() < RB; bb++)
                        if (std::find(nx.begin(), nx.end(), bb) == nx.end()) nx.push_back(bb);
                    change_to(nx, score);
                }
                seq.push_back(order[pos]);
            }
            // a first LDS change ahead of every gate that keeps the low bits
            // on lanes becomes the load layout
            if (!seq.empty() && seq[0] == -1 && changes[0].lds) {
                bool low = false;
                for (int b : changes[0].rs) low |= b < klo;
                for (int i = 0; i < klo && i < LOGW; i++) low |= changes[0].thr[i] != i;
                if (!low) {
                    rs_in = changes[0].rs;
                    thr_in = changes[0].thr;
                    seq.erase(seq.begin());
                    changes.erase(changes.begin());
                }
            }
            return;
        }
        std::vector<std::vector<int>> &succ = succ0;
        std::vector<int> &indeg = indeg0;
        std::vector<char> done(m, 0);
        int left = m;
        while (left > 0) {
            bool progress = true;
            while (progress) {
                progress = false;
                for (int a = 0; a < m; a++) {
                    if (done[a] || indeg[a]) continue;
                    const SG &g = run[order[a]];
                    if (needs_reg(g) && !in_rs(lbit[g.target])) continue;
                    seq.push_back(order[a]);
                    done[a] = 1;
                    left--;
                    for (int b : succ[a]) indeg[b]--;
                    progress = true;
                }
            }
            if (left == 0) break;
            // next layout: score target bits of the remaining non-diagonal
            // gates, ready ones first; among the RB-subsets of the best
            // scored bits, take the one under which the most gates run
            // before the next change, less the cost of reaching it
            std::vector<long> score = pair_score(indeg, &done);
            std::vector<int> nx = best_set(run, order, lbit, succ, indeg, done, score, 0, &RS, &THR);
            change_to(nx, score);
            if ((int)changes.size() > 2 * m + 8) {  // every change runs a gate: cannot happen
                bad = true;
                return;
            }
        }
    }

    void build(const std::vector<SG> &run, const std::vector<int> &order, uint64_t smask)
    {
        Sweep sw;
        memset(&sw.desc, 0, sizeof sw.desc);
        sw.tile = true;
        sw.gate = -1;
        sw.can_carry = false;
        sw.carry_z = 0;
        sw.carry_w = 0;
        int lbit[64];
        for (int q = 0; q < 64; q++) lbit[q] = -1;
        int i = 0;
        for (int q = 0; q < n; q++)
            if ((smask >> q) & 1) {
                sw.desc.spos[i] = (int8_t)q;
                lbit[q] = i++;
            }
        for (; i < 16; i++) sw.desc.spos[i] = 64;
        sw.op_begin = prog.size();
        std::vector<int> seq, RS, THR;
        std::vector<Change> changes;
        sequence(run, order, lbit, seq, changes, RS, THR);
        // load layout: register bits RS, thread bits THR
        auto reg_offsets = [&](const std::vector<int> &rs, uint64_t *out) {
            for (int p = 0; p < (1 << RB); p++) {
                uint64_t g = 0;
                for (int j = 0; j < RB; j++)
                    if ((p >> j) & 1) g |= 1ull << sw.desc.spos[rs[j]];
                out[p] = g;
            }
        };
        reg_offsets(RS, sw.desc.gr);
        for (int t = 0; t < LOGW; t++) {
            sw.desc.tin[t] = (int8_t)THR[t];
            sw.desc.tg[t] = 1ull << sw.desc.spos[THR[t]];
        }
        // fast sweeps: uncontrolled H ops leave out their 1/sqrt(2) and
        // uncontrolled diagonals their phase on |0>; both go into one global
        // factor (gre, gim) that the store applies
        size_t hdebt = 0;
        double gre = 1.0, gim = 0.0;
        Frame fr;
        // Consecutive fast sweeps: the previous one stores with only the X
        // part of its frame applied, and this one starts from the rest (its Z
        // part and phase), so the sign work at the store is done once, by
        // the last tile sweep before a plain pass or the end of the run.
        if (fast && !sweeps.empty() && sweeps.back().tile && sweeps.back().can_carry) {
            Sweep &pv = sweeps.back();
            pv.desc.fzl = 0;
            pv.desc.fzr = 0;
            pv.desc.fzg = 0;
            pv.desc.fflags &= ~(TF_SIGN | TF_ODD);
            fr.z = pv.carry_z;
            fr.w = pv.carry_w;
        }
        size_t li_next = 0;
        for (int gi : seq) {
            if (gi < 0) {
                const Change &ch = changes[li_next++];
                if (ch.lds) {
                    if (ch.rs != RS || ch.thr != THR) emit_layout(RS, ch.rs, ch.thr);
                } else {
                    for (auto [j, tb] : ch.sw) {
                        emit_swap(j, tb, RS[j], THR[tb]);
                        std::swap(RS[j], THR[tb]);
                    }
                }
                RS = ch.rs;
                THR = ch.thr;
                continue;
            }
            SG g = run[gi];
            // Pauli frame (fast sweeps): Paulis are absorbed, H / CX / CZ carry
            // the frame (updated below, once the op is out), other gates are
            // conjugated
            enum { KEEP, PUSH_H, PUSH_CX, PUSH_CZ } push = KEEP;
            const uint64_t tb = 1ull << g.target;
            const bool fa = (fr.x & tb) != 0, fb = (fr.z & tb) != 0;
            if (fast) {
                const bool unc = !g.ones && !g.zeros;
                const bool one_ctrl = !g.zeros && g.ones && !(g.ones & (g.ones - 1));
                if (absorbed(g)) {
                    if (g.k.ak == AK_X) {
                        fr.x ^= tb;
                    } else if (g.k.ak == AK_Y) {  // Y = i X Z
                        fr.w += 1 + 2 * fa;
                        fr.x ^= tb;
                        fr.z ^= tb;
                    } else {  // Z
                        fr.w += 2 * fa;
                        fr.z ^= tb;
                    }
                    frame_gates++;
                    continue;
                }
                if (unc && g.k.ak == AK_H) {
                    push = PUSH_H;
                } else if (one_ctrl && g.k.ak == AK_X) {
                    push = PUSH_CX;
                } else if (one_ctrl && g.k.ak == AK_NEG) {
                    push = PUSH_CZ;
                } else {
                    // an X in the frame on a control inverts the condition
                    const uint64_t fl = (g.ones | g.zeros) & fr.x;
                    const uint64_t o2 = (g.ones & ~fl) | (g.zeros & fl);
                    g.zeros = (g.zeros & ~fl) | (g.ones & fl);
                    g.ones = o2;
                    conj_coef(g.k, fa, fb);
                }
            }
            // fast arithmetic: rotations and phases as shears (see
            // shear_form); a controlled one beyond pi/2 runs as two halves
            bool zafter = false;
            int reps = 1;
            if (fast) {
                const bool unc = !g.ones && !g.zeros;
                if (g.k.ak == AK_RZ && unc) {  // d0 to the global factor, d1/d0 on |1>
                    const double d0r = g.k.c[0], d0i = g.k.c[1], d1r = g.k.c[2], d1i = g.k.c[3];
                    const double t = gre * d0r - gim * d0i;
                    gim = gre * d0i + gim * d0r;
                    gre = t;
                    memset(g.k.c, 0, sizeof g.k.c);
                    g.k.ak = AK_PHASE;
                    g.k.c[0] = d1r * d0r + d1i * d0i;
                    g.k.c[1] = d1i * d0r - d1r * d0i;
                }
                if (unc && (g.k.ak == AK_U || g.k.ak == AK_MAT)) {
                    double pr, pi;
                    ush_form(g.k, &pr, &pi, &zafter);
                    const double t = gre * pr - gim * pi;
                    gim = gre * pi + gim * pr;
                    gre = t;
                }
                int sign = 0;
                reps = shear_form(g.k, unc, &sign);
                if (sign == 1) fr.w += 2;     // the op ran as -G: a global -1
                if (sign == 2) zafter = true; // P ran as Z P: Z on phi (after it)
                split_ops += reps - 1;
            }
            for (int rep = 0; rep < reps; rep++) {  // (a split gate: two equal halves)
                TileOp op;
                memset(&op, 0, sizeof op);
                op.ak = g.k.ak;
                for (int j = 0; j < 8; j++) op.c[j] = g.k.c[j];
                const bool uncontrolled = !g.ones && !g.zeros;
                auto slot_of = [&](int b) {
                    for (int j = 0; j < RB; j++)
                        if (RS[j] == b) return j;
                    return -1;
                };
                uint32_t mro = 0, mrz = 0;  // controls on register-index bits (read 1 / 0)
                for (int q = 0; q < n; q++) {
                    const bool one = (g.ones >> q) & 1, zero = (g.zeros >> q) & 1;
                    if (!one && !zero) continue;
                    const int b = lbit[q];
                    if (b < 0) {
                        if (one) op.eo |= 1ull << q;
                        op.em |= 1ull << q;
                        continue;
                    }
                    const int j = slot_of(b);
                    if (j >= 0) {
                        if (one) mro |= 1u << j; else mrz |= 1u << j;
                    } else {
                        if (one) op.mlo |= 1u << b; else op.mlz |= 1u << b;
                    }
                }
                for (uint32_t p = 0; p < (1u << RB); p++)
                    if ((p & mro) == mro && !(p & mrz)) op.pm |= 1u << p;
                const bool pred = (mro | mrz | op.mlo | op.mlz) != 0;
                const uint32_t kbit = kind_bit(op.ak);
                if (!kbit) {
                    bad = true;
                    return;
                }
                op.grp = TG_GATE | kbit;
                if (fast && g.k.ak == AK_H) {
                    if (uncontrolled) hdebt++;
                    else op.grp = (op.grp & ~TG_H) | TG_HS;  // applies its own factor
                }
                if (g.diag) {
                    op.type = TOP_DIAG;
                    const int b = lbit[g.target];
                    if (b < 0) {
                        op.dsel = 2;
                        op.ebit = 1ull << g.target;
                        op.grp |= pred ? TG_DLANEP : TG_DLANE;
                    } else if (slot_of(b) >= 0) {
                        op.dsel = 0;
                        op.dpos = slot_of(b);
                        op.sel = 1u << op.dpos;
                        op.grp |= pred ? TG_DREGP : TG_DREG;
                    } else {
                        op.dsel = 1;
                        op.dpos = b;
                        op.grp |= pred ? TG_DLANEP : TG_DLANE;
                    }
                } else {
                    op.type = TOP_PAIR;
                    const int slot = slot_of(lbit[g.target]);
                    if (slot < 0) {  // sequence() switches layouts before such a gate
                        bad = true;
                        return;
                    }
                    op.sel = 1u << slot;
                    op.grp |= pred ? TG_PAIRP : TG_PAIR;
                }
                prog.push_back(op);
            }
            if (zafter) fr.z ^= tb;  // (on phi: the frame takes it on its right)
            if (push == PUSH_H) {  // H X^a Z^b = (-1)^(ab) X^b Z^a H
                fr.x = (fr.x & ~tb) | (fb ? tb : 0);
                fr.z = (fr.z & ~tb) | (fa ? tb : 0);
                fr.w += 2 * (fa && fb);
            } else if (push == PUSH_CX) {  // X_c -> X_c X_t, Z_t -> Z_c Z_t
                const uint64_t cb = g.ones;
                if (fr.x & cb) fr.x ^= tb;
                if (fb) fr.z ^= cb;
            } else if (push == PUSH_CZ) {  // X_c -> X_c Z_t, X_t -> Z_c X_t
                const uint64_t cb = g.ones;
                const bool xc = (fr.x & cb) != 0;
                fr.w += 2 * (xc && fa);
                if (xc) fr.z ^= tb;
                if (fa) fr.z ^= cb;
            }
        }
        if (fast) {
            const double f = (hdebt & 1) ? std::ldexp(0.70710678118654752440, -(int)(hdebt / 2))
                                         : std::ldexp(1.0, -(int)(hdebt / 2));
            gre *= f;
            gim *= f;
            if (gre != 1.0 || gim != 0.0) {
                sw.desc.gre = gre;
                sw.desc.gim = gim;
                sw.desc.fflags |= TF_PHASE;
            }
        }
        // store layout: the last one, unless it holds a low bit in registers
        // (then those register bits are swapped for unused high bits) or the
        // low lane bits do not hold the low local bits in order
        {
            std::vector<int> st = RS;
            uint32_t used = 0;
            for (int b : RS) used |= 1u << b;
            int nb = K - 1;
            for (int j = 0; j < RB; j++) {
                if (st[j] >= klo) continue;
                while (nb >= klo && ((used >> nb) & 1)) nb--;
                if (nb < klo) {  // too few high bits: the top RB local bits
                    for (int i = 0; i < RB; i++) st[i] = LOGW + i;
                    break;
                }
                st[j] = nb;
                used |= 1u << nb;
            }
            bool lanes_ok = true;
            for (int i = 0; i < klo && i < LOGW; i++) lanes_ok &= THR[i] == i;
            if (st != RS || !lanes_ok) emit_layout(RS, st, thread_layout(st, nullptr, 0));
            reg_offsets(st, sw.desc.gs);
            // the frame, split by where the store layout holds each bit
            if (fr.x & ~smask) {  // frames only flip qubits the sweep targets
                bad = true;
                return;
            }
            uint32_t xl = 0, zl = 0, regm = 0;
            for (int q = 0; q < n; q++) {
                if (lbit[q] < 0) continue;
                if ((fr.x >> q) & 1) xl |= 1u << lbit[q];
                if ((fr.z >> q) & 1) zl |= 1u << lbit[q];
            }
            uint64_t xg = 0;
            for (int j = 0; j < RB; j++) {
                regm |= 1u << st[j];
                if ((xl >> st[j]) & 1) xg |= 1ull << sw.desc.spos[st[j]];
            }
            for (int p = 0; p < (1 << RB); p++) sw.desc.gs[p] ^= xg;
            uint32_t fzr = 0;
            for (int p = 0; p < (1 << RB); p++) {
                uint32_t par = 0;
                for (int j = 0; j < RB; j++) par ^= ((p >> j) & 1) & ((zl >> st[j]) & 1);
                fzr |= par << p;
            }
            const int w = fr.w & 3;
            if (w & 2) fzr ^= (1u << (1 << RB)) - 1;
            sw.desc.fxl = xl & ~regm;
            sw.desc.fzl = zl & ~regm;
            sw.desc.fzr = fzr;
            sw.desc.fzg = fr.z & ~smask;
            if (sw.desc.fzl || sw.desc.fzr || sw.desc.fzg) sw.desc.fflags |= TF_SIGN;
            if (w & 1) sw.desc.fflags |= TF_SIGN | TF_ODD;
            // what a next sweep takes over if this one stores X^x phi only:
            // i^w X^x Z^z = i^(w + 2 |x & z|) Z^z X^x
            sw.can_carry = fast;
            sw.carry_z = fr.z;
            sw.carry_w = (fr.w + 2 * __builtin_popcountll(fr.x & fr.z)) & 3;
        }
        sw.op_count = prog.size() - sw.op_begin;
        sw.desc.nops = (int32_t)sw.op_count;
        sweeps.push_back(sw);
    }

    void plan(const std::vector<SG> &run)
    {
        std::vector<int> rem(run.size());
        for (size_t i = 0; i < run.size(); i++) rem[i] = (int)i;
        while (!rem.empty()) {
            uint64_t smask;
            std::vector<int> order = collect(run, rem, smask);
            if (order.size() >= 2) endgame(run, rem, smask, order);
            if (order.size() < 2) {
                // nothing to fuse: the first remaining gate runs on the plain path
                Sweep sw;
                memset(&sw, 0, sizeof sw);
                sw.tile = false;
                sw.gate = rem[0];
                sweeps.push_back(sw);
                rem.erase(rem.begin());
                continue;
            }
            build(run, order, smask);
            rem = minus(rem, order);
        }
    }
};

// Device programs: a ring of two slots, each a pinned host staging buffer, a
// device buffer (both grown as needed) and an event recorded after the last
// sweep that reads it.  A call plans on the host while the GPU may still run
// the previous call's sweeps, and waits only before it overwrites a slot
// whose sweeps have not finished: execute calls return without a device
// synchronisation, in stream order like every other entry point.
struct ProgSlot {
    TileOp *host = nullptr;
    TileOp *dev = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};
ProgSlot g_slots[2];
int g_slot_next = 0;
std::mutex g_slot_mu;

int upload_program(const std::vector<TileOp> &prog, ProgSlot **out)
{
    ProgSlot &sl = g_slots[g_slot_next];
    g_slot_next ^= 1;
    if (sl.pending) {
        SPZ_HIP(hipEventSynchronize(sl.done));
        sl.pending = false;
    }
    if (!sl.done) SPZ_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    if (prog.size() > sl.cap) {
        if (sl.dev) SPZ_HIP(hipFree(sl.dev));
        if (sl.host) SPZ_HIP(hipHostFree(sl.host));
        sl.dev = nullptr;
        sl.host = nullptr;
        sl.cap = 0;
        const size_t cap = std::max(prog.size(), (size_t)4096);
        if (hipMalloc(&sl.dev, cap * sizeof(TileOp)) != hipSuccess ||
            hipHostMalloc(&sl.host, cap * sizeof(TileOp), hipHostMallocDefault) != hipSuccess)
            return fail(SPZ_ERR_OOM, "tile program allocation failed");
        sl.cap = cap;
    }
    memcpy(sl.host, prog.data(), prog.size() * sizeof(TileOp));
    SPZ_HIP(hipMemcpyAsync(sl.dev, sl.host, prog.size() * sizeof(TileOp), hipMemcpyHostToDevice,
                           rt().stream));
    *out = &sl;
    return SPZ_OK;
}

int run_fused(spz_state_s *s, const std::vector<SG> &run, int K, bool strict)
{
    Planner pl;
    pl.n = s->n;
    pl.K = K;
    pl.LOGW = K - 4;
    pl.RB = 4;
    pl.klo = std::min(env_int("SPZ_TILE_KLO", 4), K);
    pl.improve = env_int("SPZ_TILE_SEARCH", 16);
    pl.strict = strict;
    pl.fast = !strict;
    // In-wavefront lane swaps are off unless asked for: they leave high qubits
    // on lane bits 4 and 5 at load time (four 128-byte runs per wave instead
    // of one 512-byte run), which now costs more than the LDS transposes they
    // save (30 q bench circuit: 7227 against 7190 gates/s without them).
    pl.swaps = env_int("SPZ_TILE_SWAPS", 0) != 0;
    pl.plan(run);
    if (pl.bad) return fail(SPZ_ERR_ARG, "tile planner produced a gate outside the register layout");
    if (pl.fast) {
        // The fast sweeps' global factors are scalars, and every gate in
        // between is linear: a sweep may leave its factor to a later one.
        // Only the last tile sweep of the run applies the product (or an
        // earlier one, once its magnitude leaves [2^-32, 2^32], so that the
        // stored amplitudes stay within 2^32 of the state's).
        size_t last = pl.sweeps.size();
        for (size_t i = 0; i < pl.sweeps.size(); i++)
            if (pl.sweeps[i].tile) last = i;
        double ar = 1.0, ai = 0.0;
        for (size_t i = 0; i < pl.sweeps.size(); i++) {
            if (!pl.sweeps[i].tile) continue;
            TileDesc &d = pl.sweeps[i].desc;
            if (d.fflags & TF_PHASE) {
                const double t = ar * d.gre - ai * d.gim;
                ai = ar * d.gim + ai * d.gre;
                ar = t;
            }
            const double m2 = ar * ar + ai * ai;
            if (i == last || m2 > 0x1p64 || m2 < 0x1p-64) {
                d.gre = ar;
                d.gim = ai;
                if (ar != 1.0 || ai != 0.0) d.fflags |= TF_PHASE;
                else d.fflags &= ~TF_PHASE;
                ar = 1.0;
                ai = 0.0;
            } else {
                d.fflags &= ~TF_PHASE;
            }
        }
    }
    std::lock_guard<std::mutex> lk(g_slot_mu);
    ProgSlot *slot = nullptr;
    if (!pl.prog.empty()) SPZ_TRY(upload_program(pl.prog, &slot));
    const TileOp *d_prog = slot ? slot->dev : nullptr;
    // on every exit after the upload: mark the slot busy until the stream
    // passes this point (an error return included, with the copy in flight)
    struct Release {
        ProgSlot *sl;
        ~Release()
        {
            if (!sl) return;
            if (hipEventRecord(sl->done, rt().stream) == hipSuccess) sl->pending = true;
            else (void)hipStreamSynchronize(rt().stream);
        }
    } release{slot};
    for (const Sweep &sw : pl.sweeps) {
        if (sw.tile) {
            SPZ_TRY(launch_tile(s, K, sw.desc, d_prog + sw.op_begin, pl.fast));
        } else {
            const SG &g = run[sw.gate];
            if (g.diag) SPZ_TRY(launch_diag(s, g.k, g.target, g.ones, g.zeros));
            else SPZ_TRY(launch_pair(s, g.k, g.target, g.ones, g.zeros));
        }
    }
    return SPZ_OK;
}

// 2x2 matrix (row-major, interleaved re/im) of a lowered gate.
void gate_matrix2(const Coef &k, double m[8])
{
    for (int i = 0; i < 8; i++) m[i] = 0.0;
    const double *c = k.c;
    switch (k.ak) {
    case AK_H: m[0] = c[0]; m[2] = c[0]; m[4] = c[0]; m[6] = -c[0]; break;
    case AK_X: m[2] = 1.0; m[4] = 1.0; break;
    case AK_Y: m[3] = -1.0; m[5] = 1.0; break;
    case AK_RX: m[0] = c[0]; m[3] = c[1]; m[5] = c[1]; m[6] = c[0]; break;
    case AK_RY: m[0] = c[1]; m[2] = -c[0]; m[4] = c[0]; m[6] = c[1]; break;
    case AK_U:
    case AK_MAT: for (int i = 0; i < 8; i++) m[i] = c[i]; break;
    case AK_NEG: m[0] = 1.0; m[6] = -1.0; break;
    case AK_PHASE: m[0] = 1.0; m[6] = c[0]; m[7] = c[1]; break;
    case AK_RZ: m[0] = c[0]; m[1] = c[1]; m[6] = c[2]; m[7] = c[3]; break;
    default: break;
    }
}

// Device buffer for block fragments (grown as needed, reused across calls).
void *g_dfrag = nullptr;
size_t g_dfrag_cap = 0;

// Fuse runs of gates acting on at most 5 qubits into dense 32x32 blocks.
int run_dense5(spz_state_s *s, const std::vector<SG> &run)
{
    struct Blk {
        int q[5];
        std::vector<int> gates;
        int single;
    };
    // Blocks are formed greedily over the remaining gates in program order:
    // a block takes a gate whose qubits keep it within 5 qubits and that
    // commutes with every gate the block has passed over (the tile planner's
    // rule: a passed gate blocks later gates on its non-diagonal qubits, and
    // non-diagonal action on its diagonal qubits), so a block gathers the
    // gates of a 5-qubit region across layers rather than only a consecutive
    // run.  The taken gates keep their order inside the block; a block of one
    // gate, or a gate on more than 5 qubits, runs on the plain path.
    std::vector<Blk> blocks;
    const uint64_t all = s->n >= 64 ? ~0ull : ((1ull << s->n) - 1);
    std::vector<int> rem(run.size());
    for (size_t i = 0; i < run.size(); i++) rem[i] = (int)i;
    while (!rem.empty()) {
        const SG &g0 = run[rem[0]];
        if (__builtin_popcountll((g0.nd | g0.dm) & all) > 5) {
            Blk b;
            b.single = rem[0];
            blocks.push_back(b);
            rem.erase(rem.begin());
            continue;
        }
        uint64_t Q = 0, bnd = 0, bd = 0;
        std::vector<int> take, left;
        for (size_t r = 0; r < rem.size(); r++) {
            const int idx = rem[r];
            const SG &g = run[idx];
            const uint64_t qs = (g.nd | g.dm) & all;
            const bool ok = !(g.nd & (bnd | bd)) && !(g.dm & bnd) &&
                            __builtin_popcountll(Q | qs) <= 5;
            if (ok) {
                take.push_back(idx);
                Q |= qs;
            } else {
                left.push_back(idx);
                bnd |= g.nd;
                bd |= g.dm;
                if (((bnd | bd) & all) == all) {  // nothing further can pass
                    left.insert(left.end(), rem.begin() + (long)r + 1, rem.end());
                    break;
                }
            }
        }
        rem.swap(left);
        Blk b;
        b.single = -1;
        if (take.size() == 1) {
            b.single = take[0];
        } else {
            int j = 0;
            for (int q = 0; q < s->n && j < 5; q++)
                if ((Q >> q) & 1) b.q[j++] = q;
            for (int q = 0; q < s->n && j < 5; q++)
                if (!((Q >> q) & 1)) { b.q[j++] = q; Q |= 1ull << q; }
            b.gates = take;
        }
        blocks.push_back(b);
    }
    // block matrices: U = G_last ... G_first, embedded in the 5-qubit block
    std::vector<double> f64;
    std::vector<float> f32;
    std::vector<double> U(2 * 32 * 32);
    for (const Blk &b : blocks) {
        if (b.single >= 0) continue;
        for (int i = 0; i < 32 * 32; i++) { U[2 * i] = (i / 32 == i % 32) ? 1.0 : 0.0; U[2 * i + 1] = 0.0; }
        for (int gi : b.gates) {
            const SG &g = run[gi];
            double m[8];
            gate_matrix2(g.k, m);
            int tj = -1;
            uint32_t need1 = 0, need0 = 0;
            for (int j = 0; j < 5; j++) {
                if (b.q[j] == g.target) tj = j;
                if ((g.ones >> b.q[j]) & 1) need1 |= 1u << j;
                if ((g.zeros >> b.q[j]) & 1) need0 |= 1u << j;
            }
            for (uint32_t a0 = 0; a0 < 32; a0++) {
                if ((a0 >> tj) & 1) continue;
                if ((a0 & need1) != need1 || (a0 & need0)) continue;
                const uint32_t a1 = a0 | (1u << tj);
                for (int col = 0; col < 32; col++) {
                    const double xr = U[2 * (a0 * 32 + col)], xi = U[2 * (a0 * 32 + col) + 1];
                    const double yr = U[2 * (a1 * 32 + col)], yi = U[2 * (a1 * 32 + col) + 1];
                    U[2 * (a0 * 32 + col)] = m[0] * xr - m[1] * xi + m[2] * yr - m[3] * yi;
                    U[2 * (a0 * 32 + col) + 1] = m[0] * xi + m[1] * xr + m[2] * yi + m[3] * yr;
                    U[2 * (a1 * 32 + col)] = m[4] * xr - m[5] * xi + m[6] * yr - m[7] * yi;
                    U[2 * (a1 * 32 + col) + 1] = m[4] * xi + m[5] * xr + m[6] * yi + m[7] * yr;
                }
            }
        }
        block5_fragments(U.data(), s->prec, f64, f32);
    }
    const size_t bytes = s->prec == SPZ_F64 ? f64.size() * sizeof(double) : f32.size() * sizeof(float);
    if (bytes) {
        SPZ_HIP(hipStreamSynchronize(rt().stream));  // the previous fragments may still be in use
        if (bytes > g_dfrag_cap) {
            if (g_dfrag) SPZ_HIP(hipFree(g_dfrag));
            g_dfrag = nullptr;
            if (hipMalloc(&g_dfrag, bytes) != hipSuccess)
                return fail(SPZ_ERR_OOM, "block fragment allocation failed");
            g_dfrag_cap = bytes;
        }
        const void *src = s->prec == SPZ_F64 ? (const void *)f64.data() : (const void *)f32.data();
        SPZ_HIP(hipMemcpyAsync(g_dfrag, src, bytes, hipMemcpyHostToDevice, rt().stream));
    }
    size_t bi = 0;
    const size_t esz = s->prec == SPZ_F64 ? sizeof(double) : sizeof(float);
    for (const Blk &b : blocks) {
        if (b.single >= 0) {
            const SG &g = run[b.single];
            if (g.diag) SPZ_TRY(launch_diag(s, g.k, g.target, g.ones, g.zeros));
            else SPZ_TRY(launch_pair(s, g.k, g.target, g.ones, g.zeros));
            continue;
        }
        SPZ_TRY(launch_block5(s, b.q, (const char *)g_dfrag + bi * 4096 * esz));
        bi++;
    }
    if (bytes) SPZ_HIP(hipStreamSynchronize(rt().stream));
    return SPZ_OK;
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

// The tile width and low-qubit count an execution uses (options, then the
// SPZ_TILE_K / SPZ_TILE_KLO environment, then 12 and 4).
int tile_width(const spz_exec_opts *opts, int n, int *K_out)
{
    int K = (opts && opts->tile_qubits) ? opts->tile_qubits : env_int("SPZ_TILE_K", 12);
    if (K < 10 || K > 13) return fail(SPZ_ERR_ARG, "tile_qubits must be in [10, 13]");
    if (n < K + 1) K = n - 1;
    *K_out = K;
    return SPZ_OK;
}

// Planning only (no device work): how the tile executor would split `ops`.
int plan_info(int n, int prec, const spz_op *ops, uint64_t n_ops, const spz_exec_opts *opts,
              spz_plan_info *out)
{
    if (!out || (n_ops && !ops)) return fail(SPZ_ERR_ARG, "null argument");
    if (n < 1 || n > 40) return fail(SPZ_ERR_QUBIT, "qubit count must be in [1, 40]");
    memset(out, 0, sizeof *out);
    const int mode = opts ? opts->mode : SPZ_EXEC_AUTO;
    if (mode != SPZ_EXEC_AUTO && mode != SPZ_EXEC_STRICT)
        return fail(SPZ_ERR_ARG, "plan statistics cover the tile executors (AUTO, STRICT)");
    int K;
    SPZ_TRY(tile_width(opts, n, &K));
    spz_state_s st;  // shape only: lowering reads n and the precision
    memset(&st, 0, sizeof st);
    st.n = n;
    st.prec = prec;
    st.N = 1ull << n;
    uint64_t i = 0;
    while (i < n_ops) {
        if (is_barrier(ops[i]) || K < 10) {
            out->plain_gates++;
            i++;
            continue;
        }
        std::vector<SG> run;
        uint64_t j = i;
        for (; j < n_ops && !is_barrier(ops[j]); j++) {
            SG g;
            int rc = lower(&st, ops[j], nullptr, (int)j, &g);
            if (rc == -1) continue;
            if (rc != SPZ_OK) return rc;
            run.push_back(g);
        }
        Planner pl;
        pl.n = n;
        pl.K = K;
        pl.LOGW = K - 4;
        pl.RB = 4;
        pl.klo = std::min(env_int("SPZ_TILE_KLO", 4), K);
        pl.improve = env_int("SPZ_TILE_SEARCH", 16);
        pl.strict = mode == SPZ_EXEC_STRICT;
        pl.fast = !pl.strict;
        pl.swaps = env_int("SPZ_TILE_SWAPS", 0) != 0;
        pl.plan(run);
        if (pl.bad) return fail(SPZ_ERR_ARG, "tile planner produced a gate outside the register layout");
        out->tile_gates += pl.frame_gates;
        out->tile_gates -= pl.split_ops;
        out->frame_gates += pl.frame_gates;
        for (const Sweep &sw : pl.sweeps) {
            if (!sw.tile) {
                out->plain_gates++;
                continue;
            }
            out->sweeps++;
            for (size_t o = sw.op_begin; o < sw.op_begin + sw.op_count; o++) {
                if (pl.prog[o].type != TOP_LAYOUT) out->tile_gates++;
                else if (pl.prog[o].grp & TG_LAYOUT) out->layouts++;
                else out->lane_swaps++;
            }
        }
        i = j;
    }
    return SPZ_OK;
}

int execute_ops(spz_state_s *s, const spz_op *ops, uint64_t n_ops, spz_tracker *trk,
                const spz_exec_opts *opts, uint8_t *outcomes)
{
    int mode = opts ? opts->mode : SPZ_EXEC_AUTO;
    if (mode == SPZ_EXEC_PLAIN) return execute_plain(s, ops, n_ops, trk, outcomes);
    if (mode != SPZ_EXEC_AUTO && mode != SPZ_EXEC_STRICT && mode != SPZ_EXEC_DENSE5)
        return fail(SPZ_ERR_ARG, "unknown execution mode");
    const bool dense5 = mode == SPZ_EXEC_DENSE5;
    if (dense5 && s->n < 12) return execute_plain(s, ops, n_ops, trk, outcomes);
    int K;
    SPZ_TRY(tile_width(opts, s->n, &K));
    if (K < 10) return execute_plain(s, ops, n_ops, trk, outcomes);
    const bool strict = mode == SPZ_EXEC_STRICT;
    if (outcomes)
        for (uint64_t i = 0; i < n_ops; i++) outcomes[i] = 255;
    uint64_t i = 0;
    while (i < n_ops) {
        if (is_barrier(ops[i])) {
            int rc = exec_one_plain(s, ops[i], trk, outcomes ? &outcomes[i] : nullptr);
            if (rc != SPZ_OK) {
                set_error("op " + std::to_string(i) + ": " + spz_last_error());
                return rc;
            }
            i++;
            continue;
        }
        std::vector<SG> run;
        uint64_t j = i;
        for (; j < n_ops && !is_barrier(ops[j]); j++) {
            SG g;
            int rc = lower(s, ops[j], trk, (int)j, &g);
            if (rc == -1) continue;  // no-op (unflipped noise, classical control read 0)
            if (rc != SPZ_OK) {
                set_error("op " + std::to_string(j) + ": " + spz_last_error());
                return rc;
            }
            run.push_back(g);
        }
        if (!run.empty()) SPZ_TRY(dense5 ? run_dense5(s, run) : run_fused(s, run, K, strict));
        i = j;
    }
    return SPZ_OK;
}

}  // namespace spz

extern "C" int spz_plan_stats(int n, int precision, const spz_op *ops, uint64_t n_ops,
                              const spz_exec_opts *opts, spz_plan_info *out)
{
    return spz::plan_info(n, precision, ops, n_ops, opts, out);
}

