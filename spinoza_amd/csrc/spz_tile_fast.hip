// spz_tile_fast.hip -- tile sweeps with the AUTO arithmetic (FMA-form RX and
// RY, H with its factor carried to a later gate).
#include "spz_tile_kernel.h"

namespace spz {

int launch_tile_fast(spz_state_s *s, int K, const TileDesc &d, const TileOp *d_ops)
{
    return launch_fast<true>(s, K, d, d_ops);
}

}  // namespace spz
