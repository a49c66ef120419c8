/* Per-kind loops of the CPU baseline (included by spz_oracle.c once per
 * instruction set, with ORC_TGT / ORC_NAME set; see the comment there). */
ORC_TGT static void ORC_NAME(orc_rayon_1q)(double *re, double *im, int n, int t, const orc_coef *k)
{
    switch (k->kind) {
    case ORC_H: ORC_CHUNK_LOOP(orc_pair_h(re, im, s0, s1, k)); break;
    case ORC_X: ORC_CHUNK_LOOP(orc_pair_x(re, im, s0, s1)); break;
    case ORC_Y: ORC_CHUNK_LOOP(orc_pair_y(re, im, s0, s1)); break;
    case ORC_Z: ORC_CHUNK_LOOP((re[s1] = -re[s1], im[s1] = -im[s1])); break;
    case ORC_P: ORC_CHUNK_LOOP(orc_phase_mul(re, im, s1, k->c[0], k->c[1])); break;
    case ORC_RX: ORC_CHUNK_LOOP(orc_pair_rx(re, im, s0, s1, k)); break;
    case ORC_RY: ORC_CHUNK_LOOP(orc_pair_ry(re, im, s0, s1, k)); break;
    case ORC_U: ORC_CHUNK_LOOP(orc_pair_u(re, im, s0, s1, k)); break;
    case ORC_RZ: {
        /* chunk i of 2^t amplitudes takes diag[i & 1] */
        const uint64_t cs = (uint64_t)1 << t;
        const int64_t nch = (int64_t)(((uint64_t)1 << n) >> t);
        int64_t ch;
#pragma omp parallel for ORC_PAR(n) schedule(static)
        for (ch = 0; ch < nch; ch++) {
            const double mr = k->c[2 * (ch & 1)], mi = k->c[2 * (ch & 1) + 1];
            const uint64_t base = (uint64_t)ch * cs;
            for (uint64_t i = 0; i < cs; i++) orc_diag_mul(re, im, base + i, mr, mi);
        }
        break;
    }
    default: break;
    }
}

ORC_TGT static void ORC_NAME(orc_rayon_c1)(double *re, double *im, int n, int c, int t,
                                    const orc_coef *k)
{
    switch (k->kind) {
    case ORC_H: ORC_CTRL_LOOP(orc_pair_h(re, im, s0, s1, k)); break;
    case ORC_X: ORC_CTRL_LOOP(orc_pair_x(re, im, s0, s1)); break;
    case ORC_Y: ORC_CTRL_LOOP(orc_pair_y(re, im, s0, s1)); break;
    case ORC_P: ORC_CTRL_LOOP(orc_phase_mul(re, im, s1, k->c[0], k->c[1])); break;
    case ORC_RX: ORC_CTRL_LOOP(orc_pair_rx(re, im, s0, s1, k)); break;
    case ORC_RY: ORC_CTRL_LOOP(orc_pair_ry(re, im, s0, s1, k)); break;
    case ORC_U: ORC_CTRL_LOOP(orc_pair_u(re, im, s0, s1, k)); break;
    case ORC_RZ:
        ORC_CTRL_LOOP((orc_diag_mul(re, im, s0, k->c[0], k->c[1]),
                       orc_diag_mul(re, im, s1, k->c[2], k->c[3])));
        break;
    default: break;
    }
}

