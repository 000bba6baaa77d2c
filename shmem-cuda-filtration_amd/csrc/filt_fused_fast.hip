// filt_fused_fast.hip -- instantiations of the fused step's kernel for the
// fast form (filt_fused_tile.hpp), one per tile shape and for lambda = 2
// only / general Brooks-Corey exponents.
#include "filt_fused_tile.hpp"

namespace filt {

int fused_launch_fast(const FusedGrid* g, const int* bg, long long total, int parity, bool general,
                    int tile, int order, hipStream_t s)
{
    return fz::launch_form<fz::FormFast>(g, bg, total, parity, general, tile, order, s);
}

}  // namespace filt
