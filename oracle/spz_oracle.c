/*
 * spz_oracle.c -- CPU oracle for the pair-amplitude state update of the
 * spinoza simulator (QuantumCircuit / gates::apply family).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * test-suite, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 * It is never linked into, called by or shipped with the product library
 * (spinoza_amd/), which must fail loudly when its HIP build is missing.
 *
 * What it restates (reference paths are relative to the spinoza repository):
 *   - per-gate pair arithmetic of spinoza/src/gates.rs, with the same operation
 *     order and with an explicit fma() exactly where the reference writes
 *     mul_add (gates.rs:837-838, 944-945, 1005-1016, 1256-1259);
 *   - gate coefficients (sin/cos on the host, gates.rs:749-751, 864-866,
 *     969-975, 1124-1126, 1285-1302) computed with the C library's sin/cos,
 *     which is what the reference's f64::sin_cos resolves to on Linux;
 *   - the serial (correct) semantics: the rayon Hadamard of gates.rs:593
 *     (defect: im[s0] = b1 + c1) is NOT reproduced;
 *   - multi-controlled gates with correct index masks (the reference's
 *     skip-ahead loops, gates.rs:401-446 / 807-827 / 912-928 / 1209-1229, pair
 *     wrong amplitudes unless the target is below every non-control bit);
 *   - dense block unitaries (gates.rs:1388-1487) for a block at any offset;
 *   - measurement (measurement.rs:12-92) with an explicit uniform draw;
 *   - expectation values (core.rs:198-264) and the build's Ising energy;
 *   - the build's own deterministic reduction tree and inverse-CDF sampler,
 *     restated here operation for operation so GPU results can be compared
 *     bit for bit;
 *   - at the end, the CPU baseline bench.py times (orc_rayon_*): the
 *     reference's per-kind parallel loop shapes over the same arithmetic,
 *     checked bit for bit against the checker's gate functions.
 *
 * Parity unpinned: the reference draws measurement outcomes and reservoir
 * samples from an unseeded thread_rng (measurement.rs:35-36 through
 * rand_distr ^0.4.3 Binomial(1, p); core.rs:81-112 through rand ^0.8.5), so no
 * seeded value of the reference exists to compare with.  orc_measure_qubit and
 * orc_sample restate the build's own rules (outcome 1 iff u < 1 - prob0;
 * inverse-CDF over fixed-tree block sums); they are pinned by the forced-outcome
 * known-answer test, closed forms, brute-force inverse CDF on small states and
 * a chi-squared test of the sampler, not by the reference.
 *
 * Compile with -ffp-contract=off (no implicit FMA) and without -ffast-math:
 * every rounding step below is intentional.
 *
 * Parity status of each entry point is recorded next to it ("pinned by ...").
 */
#define _GNU_SOURCE /* sincos */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Gate codes (same numbering the product's C ABI uses). */
enum {
    ORC_H = 0, ORC_X = 1, ORC_Y = 2, ORC_Z = 3, ORC_P = 4, ORC_RX = 5,
    ORC_RY = 6, ORC_RZ = 7, ORC_U = 8, ORC_SWAP = 9, ORC_BITFLIP = 10,
    ORC_M = 11
};

/* gates.rs:14 -- below this qubit count the reference runs serially. */
#define ORC_LOW_QUBIT_THRESHOLD 15

static int orc_threads = 1;

void orc_set_threads(int t) { orc_threads = t < 1 ? 1 : t; }
int orc_get_threads(void) { return orc_threads; }

#define ORC_PAR(n) if (orc_threads > 1 && (n) >= ORC_LOW_QUBIT_THRESHOLD) num_threads(orc_threads)

/* ------------------------------------------------------------------------ */
/* Gate coefficients                                                         */
/* ------------------------------------------------------------------------ */

/* Per-kind coefficient block.  Layout:
 *  H      : c[0] = FRAC_1_SQRT_2                    (math.rs:5, consts.rs:5-22)
 *  RX     : c[0] = cos(th/2), c[1] = -sin(th/2)     (gates.rs:749-751)
 *  RY     : c[0] = sin(th/2), c[1] = cos(th/2)      (gates.rs:1125-1126)
 *  RZ     : c[0..3] = d0.re, d0.im, d1.re, d1.im    (gates.rs:969-975)
 *  P      : c[0] = cos(th), c[1] = sin(th)          (gates.rs:865)
 *  U      : c[0..7] = g0.re,g0.im,g1.re,g1.im,g2.re,g2.im,g3.re,g3.im
 *                                                   (gates.rs:1286-1303)
 */
typedef struct { int kind; double c[8]; } orc_coef;

/* Sine and cosine of one angle from glibc's sincos: the reference computes
 * each gate's pair with f64::sin_cos (or sin and cos of the same angle side
 * by side), which rustc emits as LLVM sin/cos intrinsics that LLVM turns
 * into one sincos call on Linux.  sincos and separate sin/cos calls differ by
 * one ulp for some angles; the product library uses sincos too. */
static void orc_sincos(double x, double *s, double *c) { sincos(x, s, c); }

static orc_coef orc_make_coef(int kind, double th, double ph, double la)
{
    orc_coef k;
    double s, c;
    memset(&k, 0, sizeof k);
    k.kind = kind;
    switch (kind) {
    case ORC_H:
        k.c[0] = 0.70710678118654752440;   /* std::f64::consts::FRAC_1_SQRT_2 */
        break;
    case ORC_RX:
        orc_sincos(th * 0.5, &s, &c);
        k.c[0] = c;
        k.c[1] = -s;
        break;
    case ORC_RY:
        orc_sincos(th * 0.5, &s, &c);
        k.c[0] = s;
        k.c[1] = c;
        break;
    case ORC_RZ:
        orc_sincos(th * 0.5, &s, &c);
        k.c[0] = c; k.c[1] = -s; k.c[2] = c; k.c[3] = s;
        break;
    case ORC_P:
        orc_sincos(th, &s, &c);
        k.c[0] = c;
        k.c[1] = s;
        break;
    case ORC_U: {
        double st, ct, sl, cl, spl, cpl, sp, cp;
        orc_sincos(th * 0.5, &st, &ct);
        orc_sincos(la, &sl, &cl);
        orc_sincos(ph + la, &spl, &cpl);
        orc_sincos(ph, &sp, &cp);
        k.c[0] = ct;         k.c[1] = 0.0;
        k.c[2] = -cl * st;   k.c[3] = -sl * st;
        k.c[4] = cp * st;    k.c[5] = sp * st;
        k.c[6] = cpl * ct;   k.c[7] = spl * ct;
        break;
    }
    default:
        break;
    }
    return k;
}

/* Gate::to_matrix (gates.rs:95-190): row-major 2x2, interleaved re/im.
 * Used by the tests to cross-check the pair arithmetic against dense math. */
int orc_gate_matrix(int kind, double th, double ph, double la, double *m)
{
    for (int i = 0; i < 8; i++) m[i] = 0.0;
    switch (kind) {
    case ORC_H: {
        double h = 0.70710678118654752440;
        m[0] = h; m[2] = h; m[4] = h; m[6] = -h;
        return 0;
    }
    case ORC_X: m[2] = 1.0; m[4] = 1.0; return 0;
    case ORC_Y: m[3] = -1.0; m[5] = 1.0; return 0;
    case ORC_Z: m[0] = 1.0; m[6] = -1.0; return 0;
    case ORC_P: m[0] = 1.0; m[6] = cos(th); m[7] = sin(th); return 0;
    case ORC_RX: {
        double t = th / 2.0;
        m[0] = cos(t); m[3] = -sin(t); m[5] = -sin(t); m[6] = cos(t);
        return 0;
    }
    case ORC_RY: {
        double t = th / 2.0;
        m[0] = cos(t); m[2] = -sin(t); m[4] = sin(t); m[6] = cos(t);
        return 0;
    }
    case ORC_RZ: {
        double t = th / 2.0;
        m[0] = cos(t); m[1] = -sin(t); m[6] = cos(t); m[7] = sin(t);
        return 0;
    }
    case ORC_U: {
        double t = th / 2.0;
        m[0] = cos(t);
        m[2] = -cos(la) * sin(t); m[3] = -sin(la) * sin(t);
        m[4] = cos(ph) * sin(t);  m[5] = sin(ph) * sin(t);
        m[6] = cos(ph + la) * cos(t); m[7] = sin(ph + la) * cos(t);
        return 0;
    }
    default:
        return -1;
    }
}

/* ------------------------------------------------------------------------ */
/* Pair arithmetic, one function per gate kind                              */
/* ------------------------------------------------------------------------ */

/* X: swap the two amplitudes (gates.rs:329-334). */
static inline void orc_pair_x(double *re, double *im, uint64_t s0, uint64_t s1)
{
    double t = re[s0]; re[s0] = re[s1]; re[s1] = t;
    t = im[s0]; im[s0] = im[s1]; im[s1] = t;
}

/* Y: psi0' = -i psi1, psi1' = i psi0 via the swap/negate sequence of
 * y_proc_chunk (gates.rs:448-466). */
static inline void orc_pair_y(double *re, double *im, uint64_t s0, uint64_t s1)
{
    double a = re[s0], b = im[s0], c = re[s1], d = im[s1];
    re[s0] = d;
    im[s0] = -c;
    re[s1] = -b;
    im[s1] = a;
}

/* H: h_apply_strat2, serial path (gates.rs:536-563). */
static inline void orc_pair_h(double *re, double *im, uint64_t s0, uint64_t s1,
                              const orc_coef *k)
{
    double h = k->c[0];
    double a = re[s0], b = im[s0], c = re[s1], d = im[s1];
    double a1 = h * a, b1 = h * b, c1 = h * c, d1 = h * d;
    re[s0] = a1 + c1;
    im[s0] = b1 + d1;
    re[s1] = a1 - c1;
    im[s1] = b1 - d1;
}

/* RX: rx_apply_target (gates.rs:703-726). */
static inline void orc_pair_rx(double *re, double *im, uint64_t s0, uint64_t s1,
                               const orc_coef *k)
{
    double cs = k->c[0], ns = k->c[1];
    double a = re[s0], b = im[s0], c = re[s1], d = im[s1];
    re[s0] = a * cs - d * ns;
    im[s0] = b * cs + c * ns;
    re[s1] = b * -ns + c * cs;
    im[s1] = d * cs + a * ns;
}

/* RY: ry_apply_strategy2 (gates.rs:1090-1122). */
static inline void orc_pair_ry(double *re, double *im, uint64_t s0, uint64_t s1,
                               const orc_coef *k)
{
    double sn = k->c[0], cs = k->c[1];
    double a = re[s0], b = im[s0], c = re[s1], d = im[s1];
    re[s0] = a * cs - c * sn;
    im[s0] = b * cs - d * sn;
    re[s1] = a * sn + c * cs;
    im[s1] = b * sn + d * cs;
}

/* Diagonal element multiply used by RZ (gates.rs:940-946):
 *   a = c.mul_add(m.re, -d * m.im);  b = c.mul_add(m.im, d * m.re)          */
static inline void orc_diag_mul(double *re, double *im, uint64_t s, double mr, double mi)
{
    double c = re[s], d = im[s];
    re[s] = fma(c, mr, -d * mi);
    im[s] = fma(c, mi, d * mr);
}

/* P on the |1> amplitude (gates.rs:829-840):
 *   re = z_re.mul_add(cos, -z_im * sin); im = z_im.mul_add(cos, z_re * sin) */
static inline void orc_phase_mul(double *re, double *im, uint64_t s1, double cs, double sn)
{
    double zr = re[s1], zi = im[s1];
    re[s1] = fma(zr, cs, -zi * sn);
    im[s1] = fma(zi, cs, zr * sn);
}

/* U: u_apply_target (gates.rs:1231-1267), g[0].im is ignored as in the
 * reference (it is zero by construction). */
static inline void orc_pair_u(double *re, double *im, uint64_t s0, uint64_t s1,
                              const orc_coef *k)
{
    double c = re[s0], d = im[s0], m = re[s1], n = im[s1];
    double a = k->c[0];
    double kk = k->c[2], l = k->c[3];
    double q = k->c[4], r = k->c[5];
    double s = k->c[6], t = k->c[7];
    double t0 = fma(a, c, fma(kk, m, -l * n));
    double t1 = fma(a, d, fma(kk, n, l * m));
    double t2 = fma(q, c, fma(-r, d, fma(s, m, -t * n)));
    double t3 = fma(q, d, fma(r, c, fma(s, n, t * m)));
    re[s0] = t0; im[s0] = t1; re[s1] = t2; im[s1] = t3;
}

/* Apply one gate kind to the pair (s0, s1) where s0 has the target bit clear
 * and s1 = s0 | 1<<target.  Z, P and RZ are diagonal: Z and P touch s1 only
 * (gates.rs:1048-1057, 829-840), RZ scales both (gates.rs:930-965). */
static inline void orc_pair(double *re, double *im, uint64_t s0, uint64_t s1,
                            const orc_coef *k)
{
    switch (k->kind) {
    case ORC_H: orc_pair_h(re, im, s0, s1, k); break;
    case ORC_X: orc_pair_x(re, im, s0, s1); break;
    case ORC_Y: orc_pair_y(re, im, s0, s1); break;
    case ORC_Z: re[s1] = -re[s1]; im[s1] = -im[s1]; break;
    case ORC_P: orc_phase_mul(re, im, s1, k->c[0], k->c[1]); break;
    case ORC_RX: orc_pair_rx(re, im, s0, s1, k); break;
    case ORC_RY: orc_pair_ry(re, im, s0, s1, k); break;
    case ORC_RZ:
        orc_diag_mul(re, im, s0, k->c[0], k->c[1]);
        orc_diag_mul(re, im, s1, k->c[2], k->c[3]);
        break;
    case ORC_U: orc_pair_u(re, im, s0, s1, k); break;
    default: break;
    }
}

/* Visit every pair (s0, s1 = s0 | 2^target) with every bit of `ones` set and
 * every bit of `zeros` clear in s0, in the reference's decomposition: pair
 * index p in [0, N / 2^(1 + #controls)) is spread into s0 by inserting a zero
 * bit at each fixed position (target and controls, ascending), then the
 * 1-controls are set.  Without controls this is the chunked loop of
 * gates.rs:336-344 (chunks of 2^(t+1), low half paired with high half:
 * s0 = chunk * 2^(t+1) + i, contiguous runs of 2^t for each thread); with one
 * control it is the insert-bit map of gates.rs:386-388.  The reference's mask
 * scans (u_c_apply, gates.rs:1354-1362, and the multi-control arms) visit the
 * same disjoint pairs, so the order does not change the result.  Threads
 * split p into contiguous ranges at n >= LOW_QUBIT_THRESHOLD, as rayon does
 * above gates.rs:14's threshold. */
static void orc_for_pairs(double *re, double *im, int n, int target,
                          uint64_t ones, uint64_t zeros, const orc_coef *k)
{
    uint64_t tb = (uint64_t)1 << target;
    uint64_t fixed = tb | ones | zeros;
    int pos[64], np = 0;
    for (int q = 0; q < n; q++)
        if ((fixed >> q) & 1) pos[np++] = q;
    uint64_t M = (uint64_t)1 << (n - np);
    int64_t p;
#pragma omp parallel for ORC_PAR(n) schedule(static)
    for (p = 0; p < (int64_t)M; p++) {
        uint64_t s0 = (uint64_t)p;
        for (int i = 0; i < np; i++) {
            int q = pos[i];
            s0 = ((s0 >> q) << (q + 1)) | (s0 & (((uint64_t)1 << q) - 1));
        }
        s0 |= ones;
        orc_pair(re, im, s0, s0 | tb, k);
    }
}

static int orc_check_qubit(int n, int q) { return q >= 0 && q < n; }

/* gates::apply (gates.rs:215-232).  `eps` is the uniform draw that
 * BitFlipNoise takes from thread_rng (gates.rs:1366-1367): the X is applied
 * when eps <= prob.  Returns 0 on success, -1 on a bad qubit/gate. */
int orc_apply(double *re, double *im, int n, int kind, double th, double ph,
              double la, int target, int t1, double eps)
{
    if (kind == ORC_SWAP) {
        /* Gate::SWAP(t0, t1) carries its own qubits (gates.rs:224); the op
         * record stores them as (target, t1). */
        int orc_swap(double *, double *, int, int, int);
        return orc_swap(re, im, n, target, t1);
    }
    if (!orc_check_qubit(n, target)) return -1;
    if (kind == ORC_BITFLIP) {
        if (eps <= th) {
            orc_coef k = orc_make_coef(ORC_X, 0, 0, 0);
            orc_for_pairs(re, im, n, target, 0, 0, &k);
        }
        return 0;
    }
    if (kind < ORC_H || kind > ORC_U) return -1;
    orc_coef k = orc_make_coef(kind, th, ph, la);
    orc_for_pairs(re, im, n, target, 0, 0, &k);
    return 0;
}

/* swap_apply (gates.rs:1376-1386). */
int orc_swap(double *re, double *im, int n, int t0, int t1)
{
    if (!orc_check_qubit(n, t0) || !orc_check_qubit(n, t1)) return -1;
    if (t0 == t1) return 0;
    uint64_t N = (uint64_t)1 << n;
    for (uint64_t i = 0; i < N; i++) {
        if (((i >> t0) & 1) == 0 && ((i >> t1) & 1) == 1) {
            uint64_t j = i + ((uint64_t)1 << t0) - ((uint64_t)1 << t1);
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    return 0;
}

/* gates::c_apply (gates.rs:248-269).  The reference has no Z arm (todo!());
 * CZ is provided here as the Z negation restricted to control = 1, which is
 * what the product implements for Controls::Single + Z. */
int orc_c_apply(double *re, double *im, int n, int kind, double th, double ph,
                double la, int control, int target)
{
    if (!orc_check_qubit(n, target) || !orc_check_qubit(n, control) || control == target)
        return -1;
    if (kind < ORC_H || kind > ORC_U) return -1;
    orc_coef k = orc_make_coef(kind, th, ph, la);
    orc_for_pairs(re, im, n, target, (uint64_t)1 << control, 0, &k);
    return 0;
}

/* gates::mc_apply (gates.rs:290-320) with correct index masks: `ones` are
 * the controls that must read 1, `zeros` those that must read 0.  The
 * reference's own mc_apply drops zero-controls from the mask (gates.rs:300-
 * 306); callers that want that behaviour pass zeros = 0.  `ry_as_rx`
 * reproduces ry_mc_apply calling the RX arithmetic (gates.rs:1209-1229). */
int orc_mc_apply(double *re, double *im, int n, int kind, double th, double ph,
                 double la, uint64_t ones, uint64_t zeros, int target, int ry_as_rx)
{
    if (!orc_check_qubit(n, target)) return -1;
    uint64_t tb = (uint64_t)1 << target;
    if ((ones | zeros) & tb) return -1;
    if (ones & zeros) return -1;
    if (n < 64 && ((ones | zeros) >> n)) return -1;
    if (kind < ORC_H || kind > ORC_U) return -1;
    if (kind == ORC_RY && ry_as_rx) kind = ORC_RX;
    orc_coef k = orc_make_coef(kind, th, ph, la);
    orc_for_pairs(re, im, n, target, ones, zeros, &k);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Dense block unitaries                                                     */
/* ------------------------------------------------------------------------ */

/* unitary_col_vec_mul (gates.rs:1388-1424): row dot products accumulated as
 *   dot_re += a*s_re - b*s_im;  dot_im += a*s_im + b*s_re                  */
static void orc_col_vec_mul(int dim, const double *ur, const double *ui,
                            const double *vr, const double *vi,
                            double *outr, double *outi)
{
    for (int row = 0; row < dim; row++) {
        double dr = 0.0, di = 0.0;
        for (int col = 0; col < dim; col++) {
            double a = ur[row * dim + col], b = ui[row * dim + col];
            dr += a * vr[col] - b * vi[col];
            di += a * vi[col] + b * vr[col];
        }
        outr[row] = dr;
        outi[row] = di;
    }
}

/* transform_u / c_transform_u (gates.rs:1427-1487): apply a 2^m x 2^m
 * matrix to qubits [t, t+m).  The reference takes the state width from the
 * matrix (correct only for a full-state matrix at t = 0); here the state
 * width is n, which agrees with the reference wherever the reference is
 * defined.  With control c >= 0 (outside the block) only blocks whose
 * control bit is 1 are transformed. */
int orc_transform_u(double *re, double *im, int n, int m, const double *ur,
                    const double *ui, int t, int control)
{
    if (m < 1 || t < 0 || t + m > n) return -1;
    if (control >= 0 && (control >= n || (control >= t && control < t + m))) return -1;
    int dim = 1 << m;
    double *vr = malloc(sizeof(double) * dim * 4);
    if (!vr) return -1;
    double *vi = vr + dim, *orr = vi + dim, *oi = orr + dim;
    uint64_t nsuffix = (uint64_t)1 << t;
    uint64_t nprefix = (uint64_t)1 << (n - m - t);
    for (uint64_t suffix = 0; suffix < nsuffix; suffix++) {
        for (uint64_t prefix = 0; prefix < nprefix; prefix++) {
            uint64_t base = (prefix << (t + m)) + suffix;
            if (control >= 0 && !((base >> control) & 1)) continue;
            for (int j = 0; j < dim; j++) {
                uint64_t kk = base + ((uint64_t)j << t);
                vr[j] = re[kk]; vi[j] = im[kk];
            }
            orc_col_vec_mul(dim, ur, ui, vr, vi, orr, oi);
            for (int j = 0; j < dim; j++) {
                uint64_t kk = base + ((uint64_t)j << t);
                re[kk] = orr[j]; im[kk] = oi[j];
            }
        }
    }
    free(vr);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Deterministic compensated reduction (the build's fixed tree)              */
/* ------------------------------------------------------------------------ */
/* Terms x[0..M) are reduced in blocks of B = T*E = 4096 terms.  In block i,
 * "thread" t accumulates x[i*B + e*T + t] for e = 0..E-1 with a TwoSum-
 * compensated running pair (S, C); the T pairs are then combined by the
 * halving tree h = T/2, ..., 1: pair[t] = comb(pair[t], pair[t+h]).  If there
 * is more than one block, the block pairs are reduced by one more block in
 * which thread t combines pairs t, t+T, t+2T, ... in order, then the same
 * tree.  The result is S + C of the final pair.  The GPU kernels perform the
 * same floating-point operations in the same order. */

#define ORC_RT 256
#define ORC_RE 16
#define ORC_RB (ORC_RT * ORC_RE)

typedef struct { double s, c; } orc_pair2;

static inline void orc_two_sum(double a, double b, double *s, double *e)
{
    double ss = a + b;
    double bb = ss - a;
    double aa = ss - bb;
    *e = (a - aa) + (b - bb);
    *s = ss;
}

static inline orc_pair2 orc_comb(orc_pair2 x, orc_pair2 y)
{
    orc_pair2 r;
    double e;
    orc_two_sum(x.s, y.s, &r.s, &e);
    r.c = (x.c + y.c) + e;
    return r;
}

static inline void orc_acc(orc_pair2 *p, double x)
{
    double e;
    orc_two_sum(p->s, x, &p->s, &e);
    p->c = p->c + e;
}

static void orc_tree(orc_pair2 *pr)
{
    for (int h = ORC_RT / 2; h >= 1; h >>= 1)
        for (int t = 0; t < h; t++) pr[t] = orc_comb(pr[t], pr[t + h]);
}

typedef double (*orc_term_fn)(const void *ctx, uint64_t j);

/* Block partials (S, C) for every block; returns the number of blocks. */
static uint64_t orc_block_partials(orc_term_fn f, const void *ctx, uint64_t M,
                                   orc_pair2 **out)
{
    uint64_t nb = (M + ORC_RB - 1) / ORC_RB;
    if (nb == 0) nb = 1;
    orc_pair2 *bp = malloc(sizeof(orc_pair2) * nb);
    int64_t i;
#pragma omp parallel for if (orc_threads > 1 && nb > 16) num_threads(orc_threads) schedule(static)
    for (i = 0; i < (int64_t)nb; i++) {
        orc_pair2 pr[ORC_RT];
        for (int t = 0; t < ORC_RT; t++) {
            pr[t].s = 0.0; pr[t].c = 0.0;
            for (int e = 0; e < ORC_RE; e++) {
                uint64_t j = (uint64_t)i * ORC_RB + (uint64_t)e * ORC_RT + t;
                if (j < M) orc_acc(&pr[t], f(ctx, j));
            }
        }
        orc_tree(pr);
        bp[i] = pr[0];
    }
    *out = bp;
    return nb;
}

static double orc_reduce(orc_term_fn f, const void *ctx, uint64_t M)
{
    orc_pair2 *bp;
    uint64_t nb = orc_block_partials(f, ctx, M, &bp);
    orc_pair2 fin;
    if (nb == 1) {
        fin = bp[0];
    } else {
        orc_pair2 pr[ORC_RT];
        for (int t = 0; t < ORC_RT; t++) {
            pr[t].s = 0.0; pr[t].c = 0.0;
            for (uint64_t j = t; j < nb; j += ORC_RT) pr[t] = orc_comb(pr[t], bp[j]);
        }
        orc_tree(pr);
        fin = pr[0];
    }
    free(bp);
    return fin.s + fin.c;
}

typedef struct { const double *re, *im; int t; int t2; double hx; int n;
                 const double *lr, *li; } orc_ctx;

static inline uint64_t orc_insert0(uint64_t j, int t)
{
    uint64_t lo = j & (((uint64_t)1 << t) - 1);
    return ((j >> t) << (t + 1)) | lo;
}

/* |psi_a|^2 as re*re + im*im (core.rs:214 powi(2) sum, no FMA). */
static double orc_term_prob0(const void *c, uint64_t j)
{
    const orc_ctx *x = c;
    uint64_t a = orc_insert0(j, x->t);
    return x->re[a] * x->re[a] + x->im[a] * x->im[a];
}

static double orc_term_norm(const void *c, uint64_t j)
{
    const orc_ctx *x = c;
    return x->re[j] * x->re[j] + x->im[j] * x->im[j];
}

/* Re <psi | O_t psi> terms a*c + b*d (core.rs:244-258), with (c, d) the
 * amplitude of O_t psi computed by the pair arithmetic of X / Y / Z. */
static double orc_term_x(const void *c, uint64_t j)
{
    const orc_ctx *x = c;
    uint64_t p = j ^ ((uint64_t)1 << x->t);
    return x->re[j] * x->re[p] + x->im[j] * x->im[p];
}

static double orc_term_y(const void *c, uint64_t j)
{
    const orc_ctx *x = c;
    uint64_t p = j ^ ((uint64_t)1 << x->t);
    double vr, vi;
    if ((j >> x->t) & 1) { vr = -x->im[p]; vi = x->re[p]; }   /* (Y psi)_1 = i psi_0  */
    else { vr = x->im[p]; vi = -x->re[p]; }                   /* (Y psi)_0 = -i psi_1 */
    return x->re[j] * vr + x->im[j] * vi;
}

static double orc_term_z(const void *c, uint64_t j)
{
    const orc_ctx *x = c;
    double vr = x->re[j], vi = x->im[j];
    if ((j >> x->t) & 1) { vr = -vr; vi = -vi; }
    return x->re[j] * vr + x->im[j] * vi;
}

static double orc_term_zz(const void *c, uint64_t j)
{
    const orc_ctx *x = c;
    double w = x->re[j] * x->re[j] + x->im[j] * x->im[j];
    int par = (int)(((j >> x->t) ^ (j >> x->t2)) & 1);
    return par ? -w : w;
}

static double orc_term_dot(const void *c, uint64_t j)
{
    const orc_ctx *x = c;
    return x->re[j] * x->lr[j] + x->im[j] * x->li[j];
}

double orc_norm(const double *re, const double *im, int n)
{
    orc_ctx c = { re, im, 0, 0, 0.0, n, 0, 0 };
    return orc_reduce(orc_term_norm, &c, (uint64_t)1 << n);
}

/* prob0 of measurement.rs:16-29 / core.rs:202-215. */
double orc_prob0(const double *re, const double *im, int n, int target)
{
    orc_ctx c = { re, im, target, 0, 0.0, n, 0, 0 };
    return orc_reduce(orc_term_prob0, &c, (uint64_t)1 << (n - 1));
}

/* qubit_expectation_value (core.rs:198-219): 2*prob0 - 1. */
double orc_qubit_expectation_value(const double *re, const double *im, int n, int target)
{
    double p0 = orc_prob0(re, im, n, target);
    return 2.0 * p0 - 1.0;
}

/* xyz_expectation_value for one target (core.rs:222-264); obs in x,y,z. */
int orc_xyz_expectation_value(const double *re, const double *im, int n,
                              char obs, int target, double *out)
{
    if (!orc_check_qubit(n, target)) return -1;
    orc_ctx c = { re, im, target, 0, 0.0, n, 0, 0 };
    orc_term_fn f;
    if (obs == 'x') f = orc_term_x;
    else if (obs == 'y') f = orc_term_y;
    else if (obs == 'z') f = orc_term_z;
    else return -1;
    *out = orc_reduce(f, &c, (uint64_t)1 << n);
    return 0;
}

double orc_expect_zz(const double *re, const double *im, int n, int q0, int q1)
{
    orc_ctx c = { re, im, q0, q1, 0.0, n, 0, 0 };
    return orc_reduce(orc_term_zz, &c, (uint64_t)1 << n);
}

/* Re <a | b> with the fixed tree. */
double orc_dot_re(const double *ar, const double *ai, const double *br,
                  const double *bi, int n)
{
    orc_ctx c = { ar, ai, 0, 0, 0.0, n, br, bi };
    return orc_reduce(orc_term_dot, &c, (uint64_t)1 << n);
}

/* lambda = H psi for the open-boundary transverse-field Ising chain
 *   H = sum_{i<n-1} Z_i Z_{i+1} + hx * sum_i X_i.
 * Per amplitude: czz = sum_i s_i s_{i+1} (an exact integer), then
 *   lam_j = czz * psi_j + hx * (psi_{j^1} + psi_{j^2} + ... in qubit order),
 * the X sum accumulated left to right. */
void orc_ising_apply(const double *re, const double *im, int n, double hx,
                     double *lr, double *li)
{
    uint64_t N = (uint64_t)1 << n;
    int64_t jj;
#pragma omp parallel for ORC_PAR(n) schedule(static)
    for (jj = 0; jj < (int64_t)N; jj++) {
        uint64_t j = (uint64_t)jj;
        int czz = 0;
        for (int i = 0; i + 1 < n; i++) czz += (((j >> i) ^ (j >> (i + 1))) & 1) ? -1 : 1;
        double xr = 0.0, xi = 0.0;
        for (int i = 0; i < n; i++) {
            uint64_t p = j ^ ((uint64_t)1 << i);
            xr = xr + re[p];
            xi = xi + im[p];
        }
        lr[j] = (double)czz * re[j] + hx * xr;
        li[j] = (double)czz * im[j] + hx * xi;
    }
}

/* Ising energy <psi|H|psi> = Re <psi | H psi> with the fixed tree. */
double orc_ising_energy(const double *re, const double *im, int n, double hx)
{
    uint64_t N = (uint64_t)1 << n;
    double *lr = malloc(sizeof(double) * 2 * N);
    double *li = lr + N;
    orc_ising_apply(re, im, n, hx, lr, li);
    double e = orc_dot_re(re, im, lr, li, n);
    free(lr);
    return e;
}

/* ------------------------------------------------------------------------ */
/* Measurement                                                               */
/* ------------------------------------------------------------------------ */

/* measure_qubit (measurement.rs:12-92).  The outcome is `forced` when it is
 * 0 or 1 (the reference's Some(v)); otherwise it is 1 iff u < prob1 with
 * prob1 = 1 - prob0 (the reference draws Binomial(1, 1 - prob0) from an
 * unseeded thread_rng; this explicit rule is the build's definition).
 * Collapse multiplies the kept half by 1/sqrt(p) (prob.sqrt().recip()) and
 * zeroes the other half; a 1 with reset applies X (measurement.rs:87-89).
 * Returns the outcome, or -1 if the chosen outcome has probability 0 (the
 * reference would divide by zero) or the target is invalid. */
int orc_measure_qubit(double *re, double *im, int n, int target, int reset,
                      int forced, double u, double *prob0_out)
{
    if (!orc_check_qubit(n, target)) return -1;
    double prob0 = orc_prob0(re, im, n, target);
    if (prob0_out) *prob0_out = prob0;
    int val;
    if (forced == 0 || forced == 1) val = forced;
    else val = (u < 1.0 - prob0) ? 1 : 0;
    double p = val == 0 ? prob0 : 1.0 - prob0;
    if (!(p > 0.0)) return -1;
    double scale = 1.0 / sqrt(p);
    uint64_t N = (uint64_t)1 << n;
    uint64_t tb = (uint64_t)1 << target;
    for (uint64_t i = 0; i < N; i++) {
        int bit = (i & tb) != 0;
        if (bit == val) { re[i] *= scale; im[i] *= scale; }
        else { re[i] = 0.0; im[i] = 0.0; }
    }
    if (val == 1 && reset) {
        orc_coef k = orc_make_coef(ORC_X, 0, 0, 0);
        orc_for_pairs(re, im, n, target, 0, 0, &k);
    }
    return val;
}

/* ------------------------------------------------------------------------ */
/* Sampling: the build's inverse-CDF rule                                    */
/* ------------------------------------------------------------------------ */
/* w_j = re_j^2 + im_j^2.  W_b = S + C of the fixed-tree block partial of
 * block b (4096 consecutive indices).  P_0 = 0, P_{b+1} = P_b + W_b (plain
 * sequential adds), total = P_nb.  For a uniform u: x = u * total; b* is the
 * first block with P_{b+1} > x (if none, the last block with W_b > 0);
 * then acc = P_b*, and for j in block b* in order: acc = acc + w_j, the
 * sample is the first j with acc > x; if none, the last j of the block with
 * w_j > 0.  This replaces the reference's reservoir sampler
 * (core.rs:65-129), whose slots come from an unseeded RNG. */
int orc_sample(const double *re, const double *im, int n, const double *u,
               uint64_t k, uint64_t *out)
{
    uint64_t N = (uint64_t)1 << n;
    orc_ctx c = { re, im, 0, 0, 0.0, n, 0, 0 };
    orc_pair2 *bp;
    uint64_t nb = orc_block_partials(orc_term_norm, &c, N, &bp);
    double *P = malloc(sizeof(double) * (nb + 1));
    double *W = malloc(sizeof(double) * nb);
    P[0] = 0.0;
    for (uint64_t b = 0; b < nb; b++) {
        W[b] = bp[b].s + bp[b].c;
        P[b + 1] = P[b] + W[b];
    }
    double total = P[nb];
    int64_t lastnz = -1;
    for (uint64_t b = 0; b < nb; b++) if (W[b] > 0.0) lastnz = (int64_t)b;
    if (lastnz < 0) { free(P); free(W); free(bp); return -1; }
    for (uint64_t s = 0; s < k; s++) {
        double x = u[s] * total;
        /* first b with P[b+1] > x: binary search over P[1..nb] */
        uint64_t lo = 0, hi = nb;           /* answer in [lo, hi) or none */
        while (lo < hi) {
            uint64_t mid = lo + (hi - lo) / 2;
            if (P[mid + 1] > x) hi = mid; else lo = mid + 1;
        }
        uint64_t b = lo < nb ? lo : (uint64_t)lastnz;
        double acc = P[b];
        uint64_t j0 = b * ORC_RB, j1 = j0 + ORC_RB;
        if (j1 > N) j1 = N;
        int64_t pick = -1, lastw = -1;
        for (uint64_t j = j0; j < j1; j++) {
            double w = re[j] * re[j] + im[j] * im[j];
            if (w > 0.0) lastw = (int64_t)j;
            acc = acc + w;
            if (acc > x) { pick = (int64_t)j; break; }
        }
        if (pick < 0) pick = lastw;
        out[s] = (uint64_t)pick;
    }
    free(P); free(W); free(bp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Seeded inputs                                                             */
/* ------------------------------------------------------------------------ */

/* SplitMix64 (Steele, Lea, Flood 2014): the counter-based generator the
 * build uses for every seeded draw. */
static inline uint64_t orc_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Uniforms in [0, 1): (x >> 11) * 2^-53. */
void orc_uniforms(uint64_t seed, uint64_t count, double *out)
{
    uint64_t st = seed;
    for (uint64_t i = 0; i < count; i++)
        out[i] = (double)(orc_splitmix64(&st) >> 11) * 0x1.0p-53;
}

/* The Gaussian random state of the 20-qubit QFT workload: re, im i.i.d.
 * N(0,1) by Box-Muller on SplitMix64 uniforms (u1 in (0,1]), normalised by
 * 1/sqrt(norm) with norm from the fixed compensated tree. */
void orc_gaussian_state(uint64_t seed, int n, double *re, double *im)
{
    uint64_t N = (uint64_t)1 << n;
    uint64_t st = seed;
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t i = 0; i < N; i++) {
        double u1 = 1.0 - (double)(orc_splitmix64(&st) >> 11) * 0x1.0p-53;
        double u2 = (double)(orc_splitmix64(&st) >> 11) * 0x1.0p-53;
        double r = sqrt(-2.0 * log(u1));
        re[i] = r * cos(two_pi * u2);
        im[i] = r * sin(two_pi * u2);
    }
    double nrm = orc_norm(re, im, n);
    double s = 1.0 / sqrt(nrm);
    for (uint64_t i = 0; i < N; i++) { re[i] *= s; im[i] *= s; }
}

/* gen_random_state (utils.rs:168-201) with a seeded generator: uniform
 * probabilities normalised by their sum, uniform phases in [0, 2*pi). */
void orc_random_state(uint64_t seed, int n, double *re, double *im)
{
    uint64_t N = (uint64_t)1 << n;
    uint64_t st = seed;
    const double two_pi = 6.283185307179586476925286766559;
    double total = 0.0;
    for (uint64_t i = 0; i < N; i++) {
        re[i] = (double)(orc_splitmix64(&st) >> 11) * 0x1.0p-53;
        total += re[i];
    }
    double rcp = 1.0 / total;
    for (uint64_t i = 0; i < N; i++) {
        double p = re[i] * rcp;
        double a = (double)(orc_splitmix64(&st) >> 11) * 0x1.0p-53 * two_pi;
        double ps = sqrt(p);
        re[i] = ps * cos(a);
        im[i] = ps * sin(a);
    }
}

/* ------------------------------------------------------------------------ */
/* Op-list executor                                                          */
/* ------------------------------------------------------------------------ */
/* One operation of a circuit, the oracle's copy of the product's op record:
 *   kind, target, angles, swap partner, control masks and the uniform draw
 * used by M / BitFlipNoise.  `classical` marks a single control that
 * QuantumCircuit::execute turns into classical control once that qubit has
 * been measured (circuit.rs:570-578). */
typedef struct {
    int32_t kind, target;
    double theta, phi, lambda;
    int32_t t1, classical;
    uint64_t ones, zeros;
    double u;
    int32_t m; int32_t pad;
    const double *u_re, *u_im;
} orc_op;

/* QuantumCircuit::execute (circuit.rs:552-600) over an op array.  tracker[0]
 * holds the measured-qubit bits and tracker[1] their values
 * (QubitTracker, circuit.rs:122-164).  Returns 0, or -(index+1) of the op
 * that failed. */
int64_t orc_execute(double *re, double *im, int n, const orc_op *ops,
                    uint64_t n_ops, uint64_t *tracker)
{
    for (uint64_t i = 0; i < n_ops; i++) {
        const orc_op *o = &ops[i];
        int rc = 0;
        if (o->kind == 12) {  /* dense block, optionally controlled */
            int ctrl = -1;
            if (o->ones) { ctrl = __builtin_ctzll(o->ones); }
            rc = orc_transform_u(re, im, n, o->m, o->u_re, o->u_im, o->target, ctrl);
        } else if (o->kind == ORC_M) {
            if (!((tracker[0] >> o->target) & 1)) {
                int v = orc_measure_qubit(re, im, n, o->target, 1, -1, o->u, 0);
                if (v < 0) rc = -1;
                else {
                    tracker[0] |= (uint64_t)1 << o->target;
                    tracker[1] &= ~((uint64_t)1 << o->target);
                    tracker[1] |= (uint64_t)v << o->target;
                }
            }
        } else if (o->kind == ORC_SWAP) {
            rc = orc_swap(re, im, n, o->target, o->t1);
        } else if (o->ones == 0 && o->zeros == 0) {
            rc = orc_apply(re, im, n, o->kind, o->theta, o->phi, o->lambda, o->target, o->t1, o->u);
        } else {
            uint64_t ones = o->ones;
            if (o->classical && __builtin_popcountll(ones) == 1 && o->zeros == 0) {
                int c = __builtin_ctzll(ones);
                if ((tracker[0] >> c) & 1) {
                    if ((tracker[1] >> c) & 1)
                        rc = orc_apply(re, im, n, o->kind, o->theta, o->phi, o->lambda, o->target, o->t1, o->u);
                    if (rc) return -(int64_t)(i + 1);
                    continue;
                }
            }
            rc = orc_mc_apply(re, im, n, o->kind, o->theta, o->phi, o->lambda, ones, o->zeros, o->target, 0);
        }
        if (rc) return -(int64_t)(i + 1);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CPU baseline: the reference's parallel loop shapes                        */
/* ------------------------------------------------------------------------ */
/* The timed CPU path.  The checker above walks every gate through one
 * generic pair visitor; the reference instead has one loop per gate kind
 * (monomorphised Rust, compiled for AVX-512/AVX2 by `multiversion`,
 * gates.rs:206-214) in two shapes:
 *  - uncontrolled gates: rayon over the 2^(n-1-t) chunks of 2^(t+1)
 *    amplitudes, a serial inner loop pairing the low half of a chunk with
 *    its high half (x_apply gates.rs:346-374, h_apply :600-616, p_apply
 *    :864-880, z_apply :1075-1088, the RX/RY/U/Y loops); RZ walks chunks of
 *    2^t multiplying chunk i by diag[i & 1] (rz_apply_strategy1 :930-965);
 *  - single-controlled gates: rayon over the N/4 pair indices, spread into
 *    s1 by inserting 1-bits at max(c, t) then min(c, t) (x_c_apply :376-399,
 *    p_c_apply :882-910); CZ is c_apply(P(pi)) as the reference has no Z arm.
 * Below gates.rs:14's threshold (15 qubits) the loops run serially.  The
 * arithmetic is the checker's (the same static inline pair functions), so
 * results are bit-identical to orc_apply / orc_c_apply (vectorising the
 * loops changes no rounding: nothing is contracted or reassociated).  The
 * loops come in AVX-512, AVX2+FMA and baseline versions chosen at run time,
 * like the reference's multiversion attribute. */

#define ORC_CHUNK_LOOP(BODY)                                                  \
    do {                                                                      \
        const uint64_t dist = (uint64_t)1 << t;                               \
        const int64_t chunks = (int64_t)(((uint64_t)1 << (n - 1)) >> t);      \
        int64_t ch;                                                           \
        _Pragma("omp parallel for ORC_PAR(n) schedule(static)")               \
        for (ch = 0; ch < chunks; ch++) {                                     \
            const uint64_t base = ((uint64_t)(2 * ch)) << t;                  \
            for (uint64_t i = 0; i < dist; i++) {                             \
                const uint64_t s0 = base + i, s1 = s0 + dist;                 \
                (void)s0; (void)s1;                                           \
                BODY;                                                         \
            }                                                                 \
        }                                                                     \
    } while (0)

#define ORC_CTRL_LOOP(BODY)                                                   \
    do {                                                                      \
        const int64_t end = (int64_t)(((uint64_t)1 << n) >> 2);              \
        const int lo = c < t ? c : t, hi = c < t ? t : c;                     \
        const uint64_t dist = (uint64_t)1 << t;                               \
        int64_t i;                                                            \
        _Pragma("omp parallel for ORC_PAR(n) schedule(static)")               \
        for (i = 0; i < end; i++) {                                           \
            const uint64_t u = (uint64_t)i;                                   \
            const uint64_t x = u + ((uint64_t)1 << (hi - 1)) +                \
                               ((u >> (hi - 1)) << (hi - 1));                 \
            const uint64_t s1 = x + ((uint64_t)1 << lo) + ((x >> lo) << lo);  \
            const uint64_t s0 = s1 - dist;                                    \
            (void)s0;                                                         \
            BODY;                                                             \
        }                                                                     \
    } while (0)

/* The loops are compiled three times (AVX-512, AVX2 + FMA, baseline x86-64)
 * and orc_rayon_apply picks one per call from the CPU's features. */
#define ORC_TGT __attribute__((target("avx512f,avx512dq,avx512vl,avx2,fma")))
#define ORC_NAME(f) f##_avx512
#include "spz_oracle_loops.h"
#undef ORC_TGT
#undef ORC_NAME
#define ORC_TGT __attribute__((target("avx2,fma")))
#define ORC_NAME(f) f##_avx2
#include "spz_oracle_loops.h"
#undef ORC_TGT
#undef ORC_NAME
#define ORC_TGT
#define ORC_NAME(f) f##_base
#include "spz_oracle_loops.h"
#undef ORC_TGT
#undef ORC_NAME

static int orc_isa(void)
{
    static int isa = -1;
    if (isa < 0) {
        __builtin_cpu_init();
        isa = (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq") &&
               __builtin_cpu_supports("avx512vl") && __builtin_cpu_supports("fma")) ? 2
              : (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) ? 1 : 0;
    }
    return isa;
}

static void orc_rayon_1q(double *re, double *im, int n, int t, const orc_coef *k)
{
    switch (orc_isa()) {
    case 2: orc_rayon_1q_avx512(re, im, n, t, k); break;
    case 1: orc_rayon_1q_avx2(re, im, n, t, k); break;
    default: orc_rayon_1q_base(re, im, n, t, k); break;
    }
}

static void orc_rayon_c1(double *re, double *im, int n, int c, int t, const orc_coef *k)
{
    switch (orc_isa()) {
    case 2: orc_rayon_c1_avx512(re, im, n, c, t, k); break;
    case 1: orc_rayon_c1_avx2(re, im, n, c, t, k); break;
    default: orc_rayon_c1_base(re, im, n, c, t, k); break;
    }
}

/* One gate in the reference's loop shape: uncontrolled single-target gates
 * and gates with exactly one 1-control.  CZ (Z with one control) runs as the
 * reference's c_apply(P(pi)); on the checker it is the exact negation, so
 * the two differ by the residue sin(pi) ~ 1.2e-16 on the touched quarter.
 * Returns -1 for anything else (SWAP, 0-controls, several controls, M,
 * noise, dense blocks). */
int orc_rayon_apply(double *re, double *im, int n, int kind, double th, double ph, double la,
                    int target, uint64_t ones, uint64_t zeros)
{
    if (!orc_check_qubit(n, target) || kind < ORC_H || kind > ORC_U || zeros) return -1;
    if (ones == 0) {
        orc_coef k = orc_make_coef(kind, th, ph, la);
        orc_rayon_1q(re, im, n, target, &k);
        return 0;
    }
    if (__builtin_popcountll(ones) != 1) return -1;
    int c = __builtin_ctzll(ones);
    if (!orc_check_qubit(n, c) || c == target) return -1;
    orc_coef k = kind == ORC_Z ? orc_make_coef(ORC_P, 3.14159265358979323846, 0, 0)
                               : orc_make_coef(kind, th, ph, la);
    orc_rayon_c1(re, im, n, c, target, &k);
    return 0;
}

/* An op list through orc_rayon_apply (the CPU baseline's execute); returns
 * 0, or -(index+1) of the first op the loop shapes above do not cover. */
int64_t orc_rayon_execute(double *re, double *im, int n, const orc_op *ops, uint64_t n_ops)
{
    for (uint64_t i = 0; i < n_ops; i++) {
        const orc_op *o = &ops[i];
        if (o->kind > ORC_U || o->classical ||
            orc_rayon_apply(re, im, n, o->kind, o->theta, o->phi, o->lambda, o->target,
                            o->ones, o->zeros))
            return -(int64_t)(i + 1);
    }
    return 0;
}
