// slc_ens_tile_f4.hip -- batched SL tile kernels, float field storage, order 4.
#include "slc_ens_tile.h"

namespace slc {

void run_ens_tile_f4(const EnsLaunch& L, int comp, int pass, int shape, unsigned grid,
                     hipStream_t s) {
    launch_tile_comp<float, 4>(L, comp, pass, shape, grid, s);
}

}  // namespace slc
