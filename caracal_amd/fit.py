"""Force matching of the QMDFF-form box parameters on the GPU.

Reference forces come from the box force field at the reference parameters;
the fit starts from parameters perturbed by a relative amount (seeded) and
runs Adam on the relative parameters u (theta = theta_ref_scale * (1 + u)),
with the loss

    L(theta) = sum_beads |F(x; theta) - F_ref|^2 / sum_beads |F_ref|^2.

The parameter gradient comes from the hand-written adjoint through
ff.make_autograd (forces forward, d(v . dE/dx)/dtheta backward): two kernel
passes per Adam step, no finite differences.
"""
import time

import numpy as np

from . import ff as cff
from . import ffbox


def perturbed_theta(theta_ref, rel=0.10, seed=3):
    rng = np.random.default_rng(seed)
    return np.asarray(theta_ref, dtype=np.float64) * (1.0 + rel * rng.choice([-1.0, 1.0], size=len(theta_ref)))


def force_match(ff, x, theta_ref, theta0, steps=50, lr=0.02):
    """x: cuda float64 tensor (nbeads, natoms, 3).  Returns a dict with the
    loss history, the fitted parameters and the seconds per Adam step."""
    import torch
    fn = cff.make_autograd(ff)
    scale = torch.tensor(np.asarray(theta_ref, dtype=np.float64), device=x.device)
    with torch.no_grad():
        _, f_ref = fn(x, scale)
    norm = float((f_ref * f_ref).sum())
    u = torch.tensor(np.asarray(theta0, dtype=np.float64) / np.asarray(theta_ref) - 1.0, device=x.device,
                     requires_grad=True)
    opt = torch.optim.Adam([u], lr=lr)
    hist = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        opt.zero_grad()
        _, f = fn(x, scale * (1.0 + u))
        loss = ((f - f_ref) ** 2).sum() / norm
        loss.backward()
        opt.step()
        hist.append(float(loss.detach()))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    theta = (scale * (1.0 + u.detach())).cpu().numpy()
    return dict(loss=hist, theta=theta, s_per_step=dt,
                rel_param_err0=float(np.max(np.abs(np.asarray(theta0) / theta_ref - 1.0))),
                rel_param_err=float(np.max(np.abs(theta / theta_ref - 1.0))))


def box_beads(system, nbeads, seed=4, sigma=0.05):
    """Bead positions of a box: the system geometry plus Gaussian bead
    displacements (bohr)."""
    rng = np.random.default_rng(seed)
    return system["x"][None] + rng.normal(0.0, sigma, (nbeads,) + system["x"].shape)


def default_problem(nmol=1000, box_angstrom=31.0, nbeads=64, seed=2):
    s = ffbox.box_system(nmol=nmol, box_angstrom=box_angstrom, seed=seed)
    return s, box_beads(s, nbeads)
