// slc_ens_tile_f6.hip -- batched SL tile kernels, float field storage, order 6.
#include "slc_ens_tile.h"

namespace slc {

void run_ens_tile_f6(const EnsLaunch& L, int comp, int pass, int shape, unsigned grid,
                     hipStream_t s) {
    launch_tile_comp<float, 6>(L, comp, pass, shape, grid, s);
}

}  // namespace slc
