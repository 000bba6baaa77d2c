// filt_fused_fp64.hip -- instantiations of the fused step's kernel for the
// fp64 form (filt_fused_tile.hpp), one per tile shape and for lambda = 2
// only (with per-cell K and porosity and one medium, or not) / general
// Brooks-Corey exponents.
#include "filt_fused_tile.hpp"

namespace filt {

int fused_launch_fp64(const FusedGrid* g, const int* bg, long long total, int parity,
                      bool general, bool fields, int tile, int order, hipStream_t s)
{
    return fz::launch_form<fz::FormFP64>(g, bg, total, parity, general, fields, tile, order, s);
}

}  // namespace filt
