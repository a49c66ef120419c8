// spz_inplace.h -- 2x2 gate arithmetic that updates amplitudes in place.
//
// Each routine is inline assembly with tied operands: the results overwrite
// the inputs' registers, ordered so that every result is written after the
// last use of the input it replaces (one or two temporaries).  Kernels that
// keep amplitudes in registers across a loop of differently-shaped updates
// then keep them in one register set (a compiler-chosen update would leave
// results in fresh registers and copy them back on every iteration).
//
// The operations and their order are those of pair_update / diag_update
// (the reference's per-gate arithmetic, FMA exactly where it writes
// mul_add), so results are bit-identical to the plain kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace spz {

// (a, b) = psi0 (re, im), (c, d) = psi1 throughout.

// H (gates.rs:547-560): a1 = h a, c1 = h c, a = a1 + c1, c = a1 - c1; same for b, d.
__device__ __forceinline__ void ip_h(double h, double &a, double &b, double &c, double &d)
{
    double t0, t1;
    asm volatile(
        "v_mul_f64 %0, %6, %2\n\t"
        "v_mul_f64 %1, %6, %4\n\t"
        "v_add_f64 %2, %0, %1\n\t"
        "v_add_f64 %4, %0, -%1\n\t"
        "v_mul_f64 %0, %6, %3\n\t"
        "v_mul_f64 %1, %6, %5\n\t"
        "v_add_f64 %3, %0, %1\n\t"
        "v_add_f64 %5, %0, -%1"
        : "=&v"(t0), "=&v"(t1), "+v"(a), "+v"(b), "+v"(c), "+v"(d)
        : "s"(h));
}

// X (gates.rs:329-334): swap psi0 and psi1.
__device__ __forceinline__ void ip_swap(double &x, double &y)
{
    double t;
    asm volatile("v_mov_b64 %2, %0\n\tv_mov_b64 %0, %1\n\tv_mov_b64 %1, %2"
                 : "+v"(x), "+v"(y), "=&v"(t));
}

// Y (gates.rs:455-464): (a, b, c, d) <- (d, -c, -b, a); negation is a multiply
// by -1 (exact, and it flips the sign of zero like -x).
__device__ __forceinline__ void ip_y(double &a, double &b, double &c, double &d)
{
    double t;
    asm volatile(
        "v_mov_b64 %4, %0\n\t"
        "v_mov_b64 %0, %3\n\t"
        "v_mov_b64 %3, %4\n\t"
        "v_mov_b64 %4, %1\n\t"
        "v_mul_f64 %1, -1.0, %2\n\t"
        "v_mul_f64 %2, -1.0, %4"
        : "+v"(a), "+v"(b), "+v"(c), "+v"(d), "=&v"(t));
}

// RX (gates.rs:721-724): a = a cs - d ns, b = b cs + c ns, c = b (-ns) + c cs,
// d = d cs + a ns, with (cs, ns) = (cos, -sin) of theta/2.
__device__ __forceinline__ void ip_rx(double cs, double ns, double &a, double &b, double &c,
                                      double &d)
{
    double t0, t1, t2, t3;
    asm volatile(
        "v_mul_f64 %0, %4, %8\n\t"   // a cs
        "v_mul_f64 %1, %7, %9\n\t"   // d ns
        "v_mul_f64 %2, %7, %8\n\t"   // d cs
        "v_mul_f64 %3, %4, %9\n\t"   // a ns
        "v_add_f64 %4, %0, -%1\n\t"
        "v_add_f64 %7, %2, %3\n\t"
        "v_mul_f64 %0, %5, %8\n\t"   // b cs
        "v_mul_f64 %1, %6, %9\n\t"   // c ns
        "v_mul_f64 %2, %5, -%9\n\t"  // b (-ns)
        "v_mul_f64 %3, %6, %8\n\t"   // c cs
        "v_add_f64 %5, %0, %1\n\t"
        "v_add_f64 %6, %2, %3"
        : "=&v"(t0), "=&v"(t1), "=&v"(t2), "=&v"(t3), "+v"(a), "+v"(b), "+v"(c), "+v"(d)
        : "s"(cs), "s"(ns));
}

// RY (gates.rs:1116-1120): a = a cs - c sn, b = b cs - d sn, c = a sn + c cs,
// d = b sn + d cs.
__device__ __forceinline__ void ip_ry(double sn, double cs, double &a, double &b, double &c,
                                      double &d)
{
    double t0, t1, t2, t3;
    asm volatile(
        "v_mul_f64 %0, %4, %9\n\t"   // a cs
        "v_mul_f64 %1, %6, %8\n\t"   // c sn
        "v_mul_f64 %2, %4, %8\n\t"   // a sn
        "v_mul_f64 %3, %6, %9\n\t"   // c cs
        "v_add_f64 %4, %0, -%1\n\t"
        "v_add_f64 %6, %2, %3\n\t"
        "v_mul_f64 %0, %5, %9\n\t"   // b cs
        "v_mul_f64 %1, %7, %8\n\t"   // d sn
        "v_mul_f64 %2, %5, %8\n\t"   // b sn
        "v_mul_f64 %3, %7, %9\n\t"   // d cs
        "v_add_f64 %5, %0, -%1\n\t"
        "v_add_f64 %7, %2, %3"
        : "=&v"(t0), "=&v"(t1), "=&v"(t2), "=&v"(t3), "+v"(a), "+v"(b), "+v"(c), "+v"(d)
        : "s"(sn), "s"(cs));
}

// U (gates.rs:1248-1266), with (cc, dd, m, n) = (a, b, c, d):
//   a = fma(g, cc, fma(k, m, -l n))          b = fma(g, dd, fma(k, n, l m))
//   c = fma(q, cc, fma(-r, dd, fma(s, m, -t n)))
//   d = fma(q, dd, fma(r, cc, fma(s, n, t m)))
// The inner terms (all that read m and n) are formed first into four
// temporaries; then c and d overwrite m and n, and a and b overwrite cc and
// dd after their last use: 14 instructions, no register copies, and every
// expression keeps the reference's operation order.
__device__ __forceinline__ void ip_u(const double *k, double &a, double &b, double &c, double &d)
{
    double x1, x2, p1, p2;
    asm volatile(
        "v_mul_f64 %0, -%14, %7\n\t"     // -t n
        "v_fma_f64 %0, %13, %6, %0\n\t"  // x1 = fma(s, m, -t n)
        "v_mul_f64 %1, %14, %6\n\t"      // t m
        "v_fma_f64 %1, %13, %7, %1\n\t"  // x2 = fma(s, n, t m)
        "v_mul_f64 %2, -%10, %7\n\t"     // -l n
        "v_fma_f64 %2, %9, %6, %2\n\t"   // p1 = fma(k, m, -l n)
        "v_mul_f64 %3, %10, %6\n\t"      // l m
        "v_fma_f64 %3, %9, %7, %3\n\t"   // p2 = fma(k, n, l m)
        "v_fma_f64 %6, -%12, %5, %0\n\t" // fma(-r, dd, x1)
        "v_fma_f64 %6, %11, %4, %6\n\t"  // c' = fma(q, cc, .)
        "v_fma_f64 %7, %12, %4, %1\n\t"  // fma(r, cc, x2)
        "v_fma_f64 %7, %11, %5, %7\n\t"  // d' = fma(q, dd, .)
        "v_fma_f64 %4, %8, %4, %2\n\t"   // a' = fma(g, cc, p1)
        "v_fma_f64 %5, %8, %5, %3"         // b' = fma(g, dd, p2)
        : "=&v"(x1), "=&v"(x2), "=&v"(p1), "=&v"(p2), "+v"(a), "+v"(b), "+v"(c), "+v"(d)
        : "s"(k[0]), "s"(k[2]), "s"(k[3]), "s"(k[4]), "s"(k[5]), "s"(k[6]), "s"(k[7]));
}

// General 2x2 complex matrix (fast sweeps only: a gate conjugated by a Pauli
// frame).  k = u00, u01, u10, u11 as (re, im) pairs:
//   psi0' = u00 psi0 + u01 psi1,   psi1' = u10 psi0 + u11 psi1.
// The terms that read the old psi1 are formed first (four temporaries), then
// psi1 and psi0 are overwritten: 16 instructions, no register copies.
__device__ __forceinline__ void ip_mat(const double *k, double &a, double &b, double &c, double &d)
{
    double p, t, x, y;
    asm volatile(
        "v_mul_f64 %0, %12, %5\n\t"      // u10r b
        "v_fma_f64 %0, %13, %4, %0\n\t"  //  + u10i a
        "v_fma_f64 %0, %15, %6, %0\n\t"  //  + u11i c           (d' less u11r d)
        "v_mul_f64 %1, %12, %4\n\t"      // u10r a
        "v_fma_f64 %1, -%13, %5, %1\n\t" //  - u10i b
        "v_fma_f64 %1, -%15, %7, %1\n\t" //  - u11i d           (c' less u11r c)
        "v_mul_f64 %2, %10, %6\n\t"      // u01r c
        "v_fma_f64 %2, -%11, %7, %2\n\t" //  - u01i d
        "v_mul_f64 %3, %10, %7\n\t"      // u01r d
        "v_fma_f64 %3, %11, %6, %3\n\t"  //  + u01i c
        "v_fma_f64 %6, %14, %6, %1\n\t"  // c' = u11r c + .
        "v_fma_f64 %7, %14, %7, %0\n\t"  // d' = u11r d + .
        "v_fma_f64 %3, %9, %4, %3\n\t"   //  + u00i a
        "v_fma_f64 %2, -%9, %5, %2\n\t"  //  - u00i b
        "v_fma_f64 %4, %8, %4, %2\n\t"   // a' = u00r a + .
        "v_fma_f64 %5, %8, %5, %3"       // b' = u00r b + .
        : "=&v"(p), "=&v"(t), "=&v"(x), "=&v"(y), "+v"(a), "+v"(b), "+v"(c), "+v"(d)
        : "s"(k[0]), "s"(k[1]), "s"(k[2]), "s"(k[3]), "s"(k[4]), "s"(k[5]), "s"(k[6]),
          "s"(k[7]));
}

// RZ multiply in place (gates.rs:940-946): x = fma(x, mr, -y mi),
// y = fma(x, mi, y mr); y is formed first, while x still holds its input.
template <bool VCOEF>
__device__ __forceinline__ void ip_rz(double mr, double mi, double &x, double &y)
{
    double t0, t1;
    if constexpr (VCOEF)
        asm volatile(
            "v_mul_f64 %0, %3, %5\n\t"
            "v_mul_f64 %1, %3, %4\n\t"
            "v_fma_f64 %3, %2, %5, %1\n\t"
            "v_fma_f64 %2, %2, %4, -%0"
            : "=&v"(t0), "=&v"(t1), "+v"(x), "+v"(y)
            : "v"(mr), "v"(mi));
    else
        asm volatile(
            "v_mul_f64 %0, %3, %5\n\t"
            "v_mul_f64 %1, %3, %4\n\t"
            "v_fma_f64 %3, %2, %5, %1\n\t"
            "v_fma_f64 %2, %2, %4, -%0"
            : "=&v"(t0), "=&v"(t1), "+v"(x), "+v"(y)
            : "s"(mr), "s"(mi));
}

// P multiply in place (gates.rs:835-838): x = fma(x, mr, -y mi),
// y = fma(y, mr, x mi).
__device__ __forceinline__ void ip_phase(double mr, double mi, double &x, double &y)
{
    double t0, t1;
    asm volatile(
        "v_mul_f64 %0, %3, %5\n\t"
        "v_mul_f64 %1, %2, %5\n\t"
        "v_fma_f64 %3, %3, %4, %1\n\t"
        "v_fma_f64 %2, %2, %4, -%0"
        : "=&v"(t0), "=&v"(t1), "+v"(x), "+v"(y)
        : "s"(mr), "s"(mi));
}

// (x, y) <- (x sg, y sg) for a per-lane sg of +1 or -1.
__device__ __forceinline__ void ip_scale(double sg, double &x, double &y)
{
    asm volatile("v_mul_f64 %0, %0, %2\n\tv_mul_f64 %1, %1, %2" : "+v"(x), "+v"(y) : "v"(sg));
}

// ---------------------------------------------------------------------------
// Fast forms (AUTO sweeps only): fewer instructions, different rounding.
// ---------------------------------------------------------------------------

// H without its 1/sqrt(2): a = a + c, c = a - c (as a_new - 2c, one FMA, so
// no temporary); the host folds the factor into a later gate.
__device__ __forceinline__ void ip_h_unscaled(double &a, double &b, double &c, double &d)
{
    asm volatile(
        "v_add_f64 %0, %0, %2\n\t"
        "v_fma_f64 %2, %2, -2.0, %0\n\t"
        "v_add_f64 %1, %1, %3\n\t"
        "v_fma_f64 %3, %3, -2.0, %1"
        : "+v"(a), "+v"(b), "+v"(c), "+v"(d));
}

// Rotations as three shears (fast sweeps): the real pair (x, y) turned by an
// angle a with |a| <= pi/2,
//   x += k1 y;  y += k2 x;  x += k1 y      (k1 = -tan(a/2), k2 = sin a),
// three FMAs in place and no temporaries (the host keeps |a| <= pi/2, where
// |k1| <= 1, and carries the sign of a larger angle elsewhere).  NEG turns by
// -a (the coefficients negated through the operand modifiers).  Two pairs are
// interleaved so that neither waits on its own FMA chain.
template <bool NEG0, bool NEG1>
__device__ __forceinline__ void ip_rot2(double k1, double k2, double &x0, double &y0, double &x1,
                                        double &y1)
{
#define SPZ_ROT_ASM(N0, N1)                                                    \
    asm volatile("v_fma_f64 %0, " N0 "%4, %1, %0\n\t"                        \
                 "v_fma_f64 %2, " N1 "%4, %3, %2\n\t"                        \
                 "v_fma_f64 %1, " N0 "%5, %0, %1\n\t"                        \
                 "v_fma_f64 %3, " N1 "%5, %2, %3\n\t"                        \
                 "v_fma_f64 %0, " N0 "%4, %1, %0\n\t"                        \
                 "v_fma_f64 %2, " N1 "%4, %3, %2"                              \
                 : "+v"(x0), "+v"(y0), "+v"(x1), "+v"(y1)                      \
                 : "s"(k1), "s"(k2))
    if constexpr (!NEG0 && !NEG1) SPZ_ROT_ASM("", "");
    else if constexpr (NEG0 && !NEG1) SPZ_ROT_ASM("-", "");
    else if constexpr (!NEG0 && NEG1) SPZ_ROT_ASM("", "-");
    else SPZ_ROT_ASM("-", "-");
#undef SPZ_ROT_ASM
}

// One pair, coefficients per lane (VCOEF, in VGPRs) or uniform.
template <bool VCOEF>
__device__ __forceinline__ void ip_rot(double k1, double k2, double &x, double &y)
{
    if constexpr (VCOEF)
        asm volatile("v_fma_f64 %0, %2, %1, %0\n\t"
                     "v_fma_f64 %1, %3, %0, %1\n\t"
                     "v_fma_f64 %0, %2, %1, %0"
                     : "+v"(x), "+v"(y)
                     : "v"(k1), "v"(k2));
    else
        asm volatile("v_fma_f64 %0, %2, %1, %0\n\t"
                     "v_fma_f64 %1, %3, %0, %1\n\t"
                     "v_fma_f64 %0, %2, %1, %0"
                     : "+v"(x), "+v"(y)
                     : "s"(k1), "s"(k2));
}

// RX(theta) with a = theta/2: (re psi0, im psi1) turns by -a and
// (im psi0, re psi1) by +a; RY(theta): (re psi0, re psi1) and
// (im psi0, im psi1) both by +a.  k = (-tan(a/2), sin a).
__device__ __forceinline__ void ip_rx_shear(double k1, double k2, double &a, double &b, double &c,
                                            double &d)
{
    ip_rot2<true, false>(k1, k2, a, d, b, c);
}

__device__ __forceinline__ void ip_ry_shear(double k1, double k2, double &a, double &b, double &c,
                                            double &d)
{
    ip_rot2<false, false>(k1, k2, a, c, b, d);
}

// (x, y) <- (x, y) (gr + i gi): the fast sweeps' global factor at the store.
__device__ __forceinline__ void ip_cmul(double gr, double gi, double &x, double &y)
{
    double t;
    asm volatile(
        "v_mul_f64 %0, %4, %2\n\t"      // gi y
        "v_mul_f64 %2, %3, %2\n\t"      // gr y
        "v_fma_f64 %2, %4, %1, %2\n\t"  // y' = gi x + gr y
        "v_fma_f64 %1, %3, %1, -%0"       // x' = gr x - gi y
        : "=&v"(t), "+v"(x), "+v"(y)
        : "s"(gr), "s"(gi));
}

}  // namespace spz
