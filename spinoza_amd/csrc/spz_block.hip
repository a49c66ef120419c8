// spz_block.hip -- dense 5-qubit blocks applied with MFMA (gfx950).
//
// A block is a 32x32 complex matrix U on qubits b0..b4.  For every one of the
// 2^(n-5) vectors v (the 32 amplitudes that agree outside the block), out =
// U v, computed as the real product
//     [yr; yi] = [[Ur, -Ui], [Ui, Ur]] [vr; vi]      (64 x 64 times 64 x 16)
// on v_mfma_f64_16x16x4f64 (FP64 storage) or v_mfma_f32_16x16x4f32 (FP32
// storage): 4 output row blocks x 16 k-steps = 64 MFMAs per 16 vectors, 256
// flops per amplitude.
//
// A workgroup (256 threads, 4 wavefronts) owns a tile of 2^12 amplitudes: the
// block qubits, qubits 0..3 (so the tile is a union of contiguous 128-byte runs)
// and the lowest remaining qubits.  The tile is staged in LDS with coalesced
// loads; each wavefront then contracts 2 groups of 16 vectors, reading its B
// fragments from LDS and keeping the 64 A fragments of the matrix in registers,
// writes the results back into the same LDS slots, and the tile is stored.
//
// MFMA lane layout (16x16x4, measured on gfx950): lane i supplies A[i % 16][i / 16]
// and B[i / 16][i % 16] for both types; accumulator slot j of lane i holds
// D[4 (i / 16) + j][i % 16] for f32 but D[4 j + i / 16][i % 16] for f64.
#include <string.h>

#include <vector>

#include "spz_internal.h"

namespace spz {

struct alignas(8) BlockDesc {
    int8_t spos[16];   // tile qubits (ascending), 12 entries
    int8_t bpos[8];    // local positions of the 5 block bits (amplitude index bit a_j)
    int8_t vpos[8];    // local positions of the 7 vector-index bits
};

namespace {

constexpr int kBK = 12;             // tile qubits
constexpr int kBT = 256;            // threads
constexpr int kBR = (1 << kBK) / kBT;  // amplitudes per thread in the load layout

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Acc;
template <> struct Acc<double> {
    using type = v4d;
    // output row (within a 16-row block) of accumulator slot j in lane group kq
    __device__ static uint32_t row(uint32_t kq, uint32_t j) { return 4u * j + kq; }
};
template <> struct Acc<float> {
    using type = v4f;
    __device__ static uint32_t row(uint32_t kq, uint32_t j) { return 4u * kq + j; }
};

__device__ __forceinline__ v4d mfma(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t dep5(uint32_t a, const int8_t *bpos)
{
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < 5; j++)
        if ((a >> j) & 1u) r |= 1u << bpos[j];
    return r;
}

__device__ __forceinline__ uint32_t dep7(uint32_t v, const int8_t *vpos)
{
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < 7; j++)
        if ((v >> j) & 1u) r |= 1u << vpos[j];
    return r;
}

// afrag: [r (4)][s (16)][lane (64)] values of the real 64x64 matrix.
template <typename T>
__global__ void __launch_bounds__(kBT)
k_block5(T *__restrict__ re, T *__restrict__ im, const BlockDesc d, const T *__restrict__ afrag)
{
    using V = typename Acc<T>::type;
    __shared__ T lre[1 << kBK], lim[1 << kBK];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const uint64_t base = insert_zeros((uint64_t)blockIdx.x, d.spos, kBK);
    uint64_t gt = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
        if ((tid >> i) & 1u) gt |= 1ull << d.spos[i];
    // stage the tile: local index l = p * 256 + tid
#pragma unroll
    for (int p = 0; p < kBR; p++) {
        uint64_t g = base | gt;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if ((p >> j) & 1) g |= 1ull << d.spos[8 + j];
        const uint32_t l = (uint32_t)p * kBT + tid;
        lre[l] = re[g];
        lim[l] = im[g];
    }
    // the matrix fragments of this lane
    T a[4][16];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int s = 0; s < 16; s++) a[r][s] = afrag[(r * 16 + s) * 64 + lane];
    __syncthreads();

    const uint32_t col = lane & 15u, kq = lane >> 4;
#pragma unroll 1
    for (int grp = 0; grp < 2; grp++) {
        const uint32_t v = (wave * 2 + grp) * 16 + col;  // vector index 0..127
        const uint32_t vl = dep7(v, d.vpos);
        V acc[4];
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] = V{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 16; s++) {
            const uint32_t c = 4u * s + kq;  // component: 0..31 real, 32..63 imaginary
            const uint32_t l = vl | dep5(c & 31u, d.bpos);
            const T b = (c < 32) ? lre[l] : lim[l];
#pragma unroll
            for (int r = 0; r < 4; r++) acc[r] = mfma(a[r][s], b, acc[r]);
        }
        // the wavefront's own vectors only: its reads above precede these writes
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t c = 16u * r + Acc<T>::row(kq, j);
                const uint32_t l = vl | dep5(c & 31u, d.bpos);
                if (c < 32) lre[l] = acc[r][j];
                else lim[l] = acc[r][j];
            }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < kBR; p++) {
        uint64_t g = base | gt;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if ((p >> j) & 1) g |= 1ull << d.spos[8 + j];
        const uint32_t l = (uint32_t)p * kBT + tid;
        re[g] = lre[l];
        im[g] = lim[l];
    }
}

}  // namespace

// Build the per-lane A fragments of the real 64x64 form of U (32x32 complex,
// row-major interleaved re/im) into out[(r*16+s)*64 + lane].
template <typename T>
static void make_afrag(const double *U, T *out)
{
    for (int r = 0; r < 4; r++)
        for (int s = 0; s < 16; s++)
            for (int lane = 0; lane < 64; lane++) {
                const int R = 16 * r + (lane & 15);
                const int C = 4 * s + (lane >> 4);
                const int rr = R & 31, cc = C & 31;
                const double ur = U[2 * (rr * 32 + cc)], ui = U[2 * (rr * 32 + cc) + 1];
                double v;
                if (R < 32) v = C < 32 ? ur : -ui;
                else v = C < 32 ? ui : ur;
                out[(r * 16 + s) * 64 + lane] = (T)v;
            }
}

// Real 64x64 fragments of one block, appended to `out` (4096 values).
void block5_fragments(const double *U, int prec, std::vector<double> &out64,
                      std::vector<float> &out32)
{
    if (prec == SPZ_F64) {
        size_t o = out64.size();
        out64.resize(o + 4096);
        make_afrag<double>(U, out64.data() + o);
    } else {
        size_t o = out32.size();
        out32.resize(o + 4096);
        make_afrag<float>(U, out32.data() + o);
    }
}

// Apply the block whose fragments are at `dfrag` (device) on qubits q[0..4]
// (amplitude index bit j of the block = qubit q[j]).
int launch_block5(spz_state_s *s, const int *q, const void *dfrag)
{
    if (s->n < kBK) return fail(SPZ_ERR_ARG, "dense 5-qubit blocks need at least 12 qubits");
    uint64_t bmask = 0;
    for (int j = 0; j < 5; j++) {
        if (q[j] < 0 || q[j] >= s->n || ((bmask >> q[j]) & 1))
            return fail(SPZ_ERR_QUBIT, "bad block qubits");
        bmask |= 1ull << q[j];
    }
    uint64_t smask = bmask | 0xFull;
    for (int b = 0; b < s->n && __builtin_popcountll(smask) < kBK; b++) smask |= 1ull << b;
    BlockDesc d;
    memset(&d, 0, sizeof d);
    int lbit[64];
    int i = 0;
    for (int b = 0; b < s->n; b++)
        if ((smask >> b) & 1) { d.spos[i] = (int8_t)b; lbit[b] = i++; }
    for (; i < 16; i++) d.spos[i] = 64;
    for (int j = 0; j < 5; j++) d.bpos[j] = (int8_t)lbit[q[j]];
    int nv = 0;
    for (int l = 0; l < kBK; l++) {
        bool isb = false;
        for (int j = 0; j < 5; j++) isb |= d.bpos[j] == l;
        if (!isb) d.vpos[nv++] = (int8_t)l;
    }
    const uint64_t tiles = s->N >> kBK;
    const double bytes = 2.0 * (double)s->N * (s->prec == SPZ_F64 ? 16.0 : 8.0);
    LaunchTimer tm(SPZ_K_DENSE, bytes);
    if (s->prec == SPZ_F64)
        hipLaunchKernelGGL((k_block5<double>), dim3((unsigned)tiles), dim3(kBT), 0, rt().stream,
                           (double *)s->re, (double *)s->im, d, (const double *)dfrag);
    else
        hipLaunchKernelGGL((k_block5<float>), dim3((unsigned)tiles), dim3(kBT), 0, rt().stream,
                           (float *)s->re, (float *)s->im, d, (const float *)dfrag);
    SPZ_HIP(hipGetLastError());
    return SPZ_OK;
}

}  // namespace spz
