// spz_tile.hip -- tile sweeps with the reference's arithmetic (STRICT and
// the executors' exact paths), and the entry that picks an instantiation.
#include "spz_tile_kernel.h"

namespace spz {

int launch_tile_fast(spz_state_s *s, int K, const TileDesc &d, const TileOp *d_ops);

int launch_tile(spz_state_s *s, int K, const TileDesc &d, const TileOp *d_ops, bool fast)
{
    return fast ? launch_tile_fast(s, K, d, d_ops) : launch_fast<false>(s, K, d, d_ops);
}

}  // namespace spz
