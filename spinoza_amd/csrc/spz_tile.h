// spz_tile.h -- fused tile sweeps: data shared by the scheduler and the kernel.
//
// A sweep reads every amplitude once, applies a run of gates, and writes it
// back.  The state is cut into 2^(n-K) tiles of 2^K amplitudes; a tile is
// the set of indices that agree on every qubit outside the tile qubit set S
// (|S| = K, always containing the lowest qubits so that each tile is a union
// of contiguous runs).  Local index bit i of a tile is qubit S[i].
//
// A workgroup of W = 2^LOGW threads owns one tile; each thread keeps
// R = 2^(K-LOGW) amplitudes in registers.  Which local bits are "register
// bits" (the RB = K-LOGW bits of the register index) is the current layout;
// a gate whose target is a register bit is a pure in-register pair update,
// with controls and diagonal phases evaluated from each amplitude's local
// index (or, for qubits outside S, once per tile).  Layout changes go through
// LDS, or, when a register bit only trades places with lane bit 4 or 5, stay
// inside the wavefront (v_permlane16/32_swap).  The layouts a sweep loads in
// and stores from keep the low local bits
// (the lowest qubits) on the low lane bits, so lanes of a wavefront read and
// write contiguous runs; the register bits at load and store are chosen per
// sweep, so that a sweep whose first or last layout meets that rule needs no
// LDS pass to enter or leave it.
#pragma once

#include <stdint.h>

namespace spz {

enum TileOpType : int32_t { TOP_PAIR = 0, TOP_DIAG = 1, TOP_LAYOUT = 2 };

// One-hot dispatch bits of TileOp::grp: the class of the op, then its
// arithmetic kind.  The kernel tests them as independent ifs.
enum TileGroup : uint32_t {
    TG_LAYOUT = 1u << 0,
    TG_GATE = 1u << 1,    // any gate (subject to the controls outside the tile)
    TG_PAIR = 1u << 2,    // 2x2 on a register bit, no controls inside the tile
    TG_PAIRP = 1u << 3,   // 2x2 on a register bit with controls inside the tile
    TG_DREG = 1u << 4,    // diagonal, target on a register bit
    TG_DREGP = 1u << 5,
    TG_DLANE = 1u << 6,   // diagonal, target on a lane bit or outside the tile
    TG_DLANEP = 1u << 7,
    TG_H = 1u << 8, TG_X = 1u << 9, TG_Y = 1u << 10, TG_RX = 1u << 11, TG_RY = 1u << 12,
    TG_U = 1u << 13,
    // fast sweeps: an uncontrolled 2x2 as P(b) RY(2a) P(d) (AK_USH), its
    // global phase carried to the sweep's global factor
    TG_USH = 1u << 14,
    TG_NEG = 1u << 15, TG_PHASE = 1u << 16,
    TG_RZ = 1u << 17,
    // in-wavefront layout changes: register-index bit `sel` trades places
    // with lane bit 4 (v_permlane16_swap) or lane bit 5 (v_permlane32_swap)
    TG_SW16 = 1u << 18, TG_SW32 = 1u << 19,
    // fast sweeps (AUTO): an H scaled by c[0] (a controlled H applies its own
    // 1/sqrt(2)); plain TG_H then adds and subtracts only (the factor of an
    // uncontrolled H goes to the sweep's global factor)
    TG_HS = 1u << 20,
    // fast sweeps: a general 2x2 complex matrix (c[0..7], row-major, re/im
    // interleaved), for gates conjugated by the sweep's Pauli frame
    TG_MAT = 1u << 21,
};
constexpr uint32_t TG_ANYLAYOUT = TG_LAYOUT | TG_SW16 | TG_SW32;

// LDS bank swizzle for 8-byte elements of a tile: XOR the low 4 bits of the
// local index with two higher nibbles so that power-of-two strides across
// lanes spread over the banks (a bijection inside every aligned block of 16).
// swz is linear over XOR, and the lane part and the register part of a local
// index have disjoint bits, so swz(lane | reg) = swz(lane) ^ swz(reg).
__host__ __device__ __forceinline__ uint32_t tile_swz(uint32_t l)
{
    return l ^ ((l >> 4) & 15u) ^ ((l >> 8) & 15u);
}

// A change of register layout.  The register part of an LDS byte offset is
// linear in the register index, so the host hands over its value on each
// register-index bit (rob/rnb) and the kernel forms the 2^RB offsets with one
// scalar XOR each.
struct TileLayout {
    int8_t newthr[16];  // local bit of thread-index bit i after the change
    uint32_t rob[4];    // 8 * tile_swz(1 << local bit of register-index bit j) before ...
    uint32_t rnb[4];    // ... and after the change
    int8_t oldreg[8];   // local bit of register-index bit j before ...
    int8_t newreg[8];   // ... and after the change (host bookkeeping)
};
constexpr int kTileLayoutRB = 4;  // register-index bits a TileLayout describes

// 128 bytes, fetched whole into scalar registers by each wavefront.
struct TileOp {
    int32_t type;     // TileOpType
    int32_t ak;       // Arith kind
    uint32_t sel;     // one-hot register-index bit of the target (PAIR, DIAG with dsel 0)
    int32_t dsel;     // DIAG target: 0 register-index bit dpos, 1 local bit dpos of the
                      // lane part, 2 outside S (global bit ebit)
    int32_t dpos;
    uint32_t pm;        // bit p: register p meets the controls on register-index bits
    uint32_t pad;
    uint32_t mlo, mlz;  // controls on local bits held by the lane part (per lane);
                        // TG_SW16/TG_SW32: the local bit leaving the lane part
                        // (mlo, as a mask) and the one entering it (mlz)
    uint32_t grp;       // TileGroup bits
    uint64_t eo, em;    // controls outside S: the op runs on a tile iff (base & em) == eo
                        // (a layout change has em = 0, eo = 1 and never runs as a gate)
    uint64_t ebit;      // DIAG with the target outside S: its global bit
    union {
        double c[8];    // PAIR / DIAG coefficients
        TileLayout lay; // LAYOUT
    };
};
static_assert(sizeof(TileOp) == 128, "TileOp must stay 128 bytes");

constexpr int kTileMaxK = 14;
constexpr int kTileMaxR = 32;

// Pauli frame and global factor of a fast sweep.  The program leaves phi in
// registers, and the state to store is psi = w X^x Z^z (g phi) (w = i^k; X^x
// flips the qubits of x, Z^z negates amplitudes whose index has odd parity on
// z; g = gre + i gim when TF_PHASE is set, else 1):
//   psi[j ^ x] = w (-1)^popcount(j & z) g phi[j].
// The host splits x and z by where the store layout holds each bit: x on
// register bits is folded into gs[], x on lane bits is fxl; z on lane bits is
// fzl, on register bits the per-register parity mask fzr (bit p), outside the
// tile fzg (global bits).  The real sign of w is folded into fzr; an
// imaginary w (k odd) stores (-im, re).
enum TileFrameFlags : uint32_t { TF_SIGN = 1u, TF_ODD = 2u, TF_PHASE = 4u };

struct TileDesc {
    int8_t spos[16];          // qubit of local bit i (ascending), K entries
    int8_t tin[16];           // local bit of thread-index bit i in the load layout
    uint64_t gr[kTileMaxR];   // global offset of register index p in the load layout
    uint64_t gs[kTileMaxR];   // ... and in the store layout (its lane part follows the
                              // program's last layout change); in amplitudes on the
                              // host, in bytes in the kernel (the launch scales them)
    int32_t nops;
    uint32_t fflags;          // TileFrameFlags (0: no sign or phase at the store)
    uint32_t fxl, fzl, fzr;   // Pauli frame (fast sweeps; zero otherwise)
    uint32_t pad;
    uint64_t fzg;
    double gre, gim;          // fast sweeps, TF_PHASE: a global factor applied before
                              // the frame (the phases the diagonals and general 2x2
                              // ops left out and the 1/sqrt(2) of the plain H, of
                              // this sweep and of the earlier sweeps of the same
                              // run that left theirs to it)
    uint64_t tg[16];          // global bit of thread-index bit i in the load layout
                              // (1 << spos[tin[i]]: uniform, no per-lane table reads)
};

// fast: the AUTO arithmetic (rotations and phases as shears, H and the |0>
// phases of diagonals carried to the store, the Pauli frame; see the kernel);
// otherwise every gate is the reference's operation sequence.
int launch_tile(spz_state_s *s, int K, const TileDesc &d, const TileOp *d_ops, bool fast);

}  // namespace spz
