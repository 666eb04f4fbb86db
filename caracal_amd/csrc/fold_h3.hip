// Folded step kernels of the H3 surface (the kernels are in fold.hpp).
// Compiled without machine-level loop-invariant code motion, as every
// translation unit with binary32 H3 step kernels (step_common.hpp, SurfH3).
#include "fold.hpp"

namespace cam {

step_fn fold_kernel_h3(int precision, int nbeads, int ensemble, int nbuf, int *tpb)
{
    if (precision == CAM_PREC_FP64)
        return nbuf == 2 ? fold::pick_fold_w<SurfH3, double, 1, 2>(nbeads, ensemble, tpb)
                         : fold::pick_fold_w<SurfH3, double, 1, 1>(nbeads, ensemble, tpb);
    return nbuf == 2 ? fold::pick_fold_w<SurfH3, float, 1, 2>(nbeads, ensemble, tpb)
                     : fold::pick_fold_w<SurfH3, float, 1, 1>(nbeads, ensemble, tpb);
}

}  // namespace cam
