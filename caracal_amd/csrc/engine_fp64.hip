// The engine's step kernels of the surfaces evaluated in FP64 only: OH + H2,
// BrH2 and O3 (step_kernels.hpp).
#include "step_kernels.hpp"

namespace cam {

bool pick_kernels_fp64(int surface, int nb, int ens, int ens_nm, int wpe, bool mfma, step_fn &st, force_fn &fo,
                       step_fn &nm, force_fn &fs)
{
    if (surface == CAM_SURFACE_OH3) return pick_nb<SurfOH3, double>(nb, ens, ens_nm, wpe, mfma, st, fo, nm, fs);
    if (surface == CAM_SURFACE_BRH2) return pick_nb<SurfBrH2, double>(nb, ens, ens_nm, wpe, mfma, st, fo, nm, fs);
    if (surface == CAM_SURFACE_O3) return pick_nb<SurfO3, double>(nb, ens, ens_nm, wpe, mfma, st, fo, nm, fs);
    return false;
}

}  // namespace cam
