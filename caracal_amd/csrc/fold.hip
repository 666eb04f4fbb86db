// Folded step kernels of the OH3, BrH2 and O3 surfaces, the dispatch over all folded
// kernels and the folded transform matrix.  The kernels themselves are in
// fold.hpp; the CH5 and H3 ones are built in fold_ch4h.hip and fold_h3.hip,
// with other code generation options.
#include <cmath>

#include "fold.hpp"

namespace cam {

step_fn fold_kernel_ch4h(int precision, int nbeads, int ensemble, int wpe, int *tpb);
step_fn fold_kernel_h3(int precision, int nbeads, int ensemble, int nbuf, int *tpb);

template <class S>
step_fn fold_fp64(int nbeads, int ensemble, int nbuf, int *tpb)
{
    return nbuf == 2 ? fold::pick_fold_w<S, double, 1, 2>(nbeads, ensemble, tpb)
                     : fold::pick_fold_w<S, double, 1, 1>(nbeads, ensemble, tpb);
}

step_fn fold_step_kernel_for(int surface, int precision, int nbeads, int ensemble, int wpe, int nbuf, int *tpb)
{
    if (ensemble != CAM_ENS_NVE && ensemble != CAM_ENS_CHILD) return nullptr;
    const bool f64 = precision == CAM_PREC_FP64;
    if (surface == CAM_SURFACE_H3) return fold_kernel_h3(precision, nbeads, ensemble, nbuf, tpb);
    if (surface == CAM_SURFACE_OH3) return f64 ? fold_fp64<SurfOH3>(nbeads, ensemble, nbuf, tpb) : nullptr;
    if (surface == CAM_SURFACE_BRH2) return f64 ? fold_fp64<SurfBrH2>(nbeads, ensemble, nbuf, tpb) : nullptr;
    if (surface == CAM_SURFACE_O3) return f64 ? fold_fp64<SurfO3>(nbeads, ensemble, nbuf, tpb) : nullptr;
    return fold_kernel_ch4h(precision, nbeads, ensemble, wpe, tpb);
}

void fold_matrix(int nbeads, double *m)
{
    // the entries of the reference's cosine matrix (cos of 2 pi (r c mod N) / N,
    // over sqrt(N)), doubled for the columns that stand for two beads
    const int h = nbeads / 2;
    const double f = std::sqrt(1.0 / nbeads), pi = 3.141592653589793;
    for (int r = 0; r < h; ++r)
        for (int c = 0; c < h; ++c) {
            const long mm = ((long)r * c) % nbeads;
            m[(size_t)r * h + c] = (c == 0 ? 1.0 : 2.0) * (std::cos(2.0 * pi * (double)mm / nbeads) * f);
        }
}

}  // namespace cam
