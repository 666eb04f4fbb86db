"""Force matching of the QMDFF-form box parameters on the GPU.

Reference forces come from the box force field at the reference parameters;
the fit starts from parameters perturbed by a relative amount (seeded) and
minimises, over the relative parameters u (theta = theta_ref_scale * (1 + u)),

    L(theta) = sum_beads |F(x; theta) - F_ref|^2 / sum_beads |F_ref|^2

with L-BFGS (strong Wolfe line search).  The parameter gradient comes from
the hand-written adjoint through ff.make_autograd (forces forward,
d(v . dE/dx)/dtheta backward): two kernel passes per evaluation, no finite
differences.

Not every parameter is fixed by forces: the Gauss-Newton matrix J^T J of the
force residual has three eigenvalues some 1e-9 of its largest, along the
three dispersion coefficients C6 (theta[7:10]; their forces are a few 1e-5
of the total), and four more near 1e-4 of it (the Morse De/a split, the
angle and charge terms).  A per-parameter normalised method (Adam) wanders
along those flat directions while the loss falls; L-BFGS builds the
curvature, recovers the seven force-determined parameters and leaves the
C6 where it started.
"""
import time

import numpy as np

from . import ff as cff
from . import ffbox

# parameters the forces determine (De, a, r0, k, theta0, q_O, q_H) and the
# dispersion coefficients C6_OO, C6_OH, C6_HH they do not
FORCE_DETERMINED = np.arange(7)
DISPERSION = np.arange(7, 10)


def perturbed_theta(theta_ref, rel=0.10, seed=3):
    rng = np.random.default_rng(seed)
    return np.asarray(theta_ref, dtype=np.float64) * (1.0 + rel * rng.choice([-1.0, 1.0], size=len(theta_ref)))


def force_match(ff, x, theta_ref, theta0, steps=50, history=20):
    """x: cuda float64 tensor (nbeads, natoms, 3).  Up to `steps` L-BFGS
    iterations.  Returns a dict with the loss after each evaluation, the
    fitted parameters, the relative parameter errors (all, the force-determined
    seven and the dispersion three, at the start and the end), the number of
    iterations and evaluations and the seconds per evaluation."""
    import torch
    fn = cff.make_autograd(ff)
    scale = torch.tensor(np.asarray(theta_ref, dtype=np.float64), device=x.device)
    with torch.no_grad():
        _, f_ref = fn(x, scale)
    norm = float((f_ref * f_ref).sum())
    u = torch.tensor(np.asarray(theta0, dtype=np.float64) / np.asarray(theta_ref) - 1.0, device=x.device,
                     requires_grad=True)
    opt = torch.optim.LBFGS([u], lr=1.0, max_iter=steps, max_eval=3 * steps, history_size=history,
                            tolerance_grad=1e-14, tolerance_change=1e-18, line_search_fn="strong_wolfe")
    hist = []

    def closure():
        opt.zero_grad()
        _, f = fn(x, scale * (1.0 + u))
        loss = ((f - f_ref) ** 2).sum() / norm
        loss.backward()
        hist.append(float(loss.detach()))
        return loss

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    opt.step(closure)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(1, len(hist))
    theta = (scale * (1.0 + u.detach())).cpu().numpy()
    err0 = np.abs(np.asarray(theta0) / theta_ref - 1.0)
    err = np.abs(theta / theta_ref - 1.0)
    return dict(loss=hist, theta=theta, s_per_eval=dt, iterations=int(opt.state[opt._params[0]]["n_iter"]),
                evaluations=len(hist), rel_param_err0=float(err0.max()), rel_param_err=float(err.max()),
                rel_err_forces0=float(err0[FORCE_DETERMINED].max()), rel_err_forces=float(err[FORCE_DETERMINED].max()),
                rel_err_dispersion0=float(err0[DISPERSION].max()), rel_err_dispersion=float(err[DISPERSION].max()))


def box_beads(system, nbeads, seed=4, sigma=0.05):
    """Bead positions of a box: the system geometry plus Gaussian bead
    displacements (bohr)."""
    rng = np.random.default_rng(seed)
    return system["x"][None] + rng.normal(0.0, sigma, (nbeads,) + system["x"].shape)


def default_problem(nmol=1000, box_angstrom=31.0, nbeads=64, seed=2):
    s = ffbox.box_system(nmol=nmol, box_angstrom=box_angstrom, seed=seed)
    return s, box_beads(s, nbeads)
