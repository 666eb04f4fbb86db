// The engine's step kernels of the CH5 surface (step_kernels.hpp), FP64 and
// mixed precision.  Compiled without machine-level loop-invariant code motion
// (the unfolded FP64 steps at two waves per SIMD need it).
#include "step_kernels.hpp"

namespace cam {

bool pick_kernels_ch4h(int precision, int nb, int ens, int ens_nm, int wpe, bool mfma, step_fn &st, force_fn &fo,
                       step_fn &nm, force_fn &fs)
{
    return precision == CAM_PREC_FP64 ? pick_nb<SurfCH4H, double>(nb, ens, ens_nm, wpe, mfma, st, fo, nm, fs)
                                      : pick_nb<SurfCH4H, float>(nb, ens, ens_nm, wpe, mfma, st, fo, nm, fs);
}

}  // namespace cam
