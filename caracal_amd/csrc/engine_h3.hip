// The engine's step kernels of the H3 surface (step_kernels.hpp), FP64 and
// mixed precision.  Compiled without machine-level loop-invariant code
// motion, as every translation unit with binary32 H3 step kernels.
#include "step_kernels.hpp"

namespace cam {

bool pick_kernels_h3(int precision, int nb, int ens, int ens_nm, int wpe, bool mfma, step_fn &st, force_fn &fo,
                     step_fn &nm, force_fn &fs)
{
    return precision == CAM_PREC_FP64 ? pick_nb<SurfH3, double>(nb, ens, ens_nm, wpe, mfma, st, fo, nm, fs)
                                      : pick_nb<SurfH3, float>(nb, ens, ens_nm, wpe, mfma, st, fo, nm, fs);
}

}  // namespace cam
