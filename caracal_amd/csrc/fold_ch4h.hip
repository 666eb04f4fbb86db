// Folded step kernels of the CH5 surface (the kernels are in fold.hpp).
// This file is compiled without machine-level loop-invariant code motion: in
// the folded step loop that pass hoists the surface's constants and lane
// indices out of the loop and holds them in registers across it, and the
// FP64 kernel then needs more than 256 registers and runs one wave per SIMD
// instead of two.
#include "fold.hpp"

namespace cam {

// A second register budget (two waves per SIMD) exists for FP64 only; the
// binary32 kernels fit three waves within the first.
step_fn fold_kernel_ch4h(int precision, int nbeads, int ensemble, int wpe, int *tpb)
{
    if (precision != CAM_PREC_FP64) return fold::pick_fold_w<SurfCH4H, float, 1, 2>(nbeads, ensemble, tpb);
    if (wpe == 2) return fold::pick_fold_w<SurfCH4H, double, 2, 2>(nbeads, ensemble, tpb);
    return fold::pick_fold_w<SurfCH4H, double, 1, 2>(nbeads, ensemble, tpb);
}

}  // namespace cam
