"""Synthetic inputs of the RPMD workloads (seeded; no data files).

h3_nve        3 H atoms (1.008 amu) on BKMP2, collinear at R = 1.757 bohr
              (examples/calc_rate/h+h2/ts.xyz), bead positions + N(0, 0.01^2)
              bohr, momenta Maxwell-Boltzmann at the bead temperature, seed 0.
ch4h_children CH4 + H recrossing children on the CH5 surface: saddle
              geometry + N(0, 0.02^2) per trajectory, centroid constrained to
              xi = 0 (dividing surface xi_ideal = 1), momenta at 500 K.
oh3_nve       OH + H2 (O 15.999, H 1.008 amu) on the Schatz-Elgersma surface at
              its saddle, bead positions + N(0, 0.01^2) bohr, momenta at 300 K.
brh2_nve      H + HBr (H 1.008, Br 79.904 amu) on the BrH2 DIM-3C surface at
              its collinear saddle, bead positions + N(0, 0.01^2) bohr, momenta
              at 300 K.
o3_nve        O + O2 (O 15.999 amu) on the Varga-Paukku-Truhlar O3 surface at
              the C2v saddle of the exchange, bead positions + N(0, 0.01^2)
              bohr, momenta at 300 K.
ragged        H3 / CH4H half-and-half with nbeads in {4,8,16,32,64} and
              T in U[200, 1000] K, seed 1 (bucketed by surface and nbeads).
"""
import numpy as np

from .engine import Reaction
from .units import FS_TO_AU, MASS_C_REF, MASS_H_REF, beta_of, mass_au

H3_TS = np.array([[0.0, 0.0, -1.757], [0.0, 0.0, 0.0], [0.0, 0.0, 1.757]])

# First-order saddle of the CH5 surface (bohr; atom order h4, C, h1, h2, h3, Hb),
# found by a partitioned rational-function Newton search on the oracle surface
# (tests/golden/make_ch4h_saddle.py): one imaginary mode, |dV/dq| < 1e-12.
CH4H_TS = np.array([
    [-1.140933199706927, -1.1409331997068821, 1.2740139038988854],
    [0.00838657829034074, 0.00838657829033506, 0.0083865782904126],
    [1.5253054531799726, 1.5253054531798804, 1.525305453180047],
    [1.2740139038988167, -1.1409331997069112, -1.1409331997068224],
    [-1.1409331997068484, 1.2740139038988751, -1.1409331997068088],
    [2.4775365649116083, 2.477536564911461, 2.477536564911683],
])

# First-order saddle of the OH + H2 surface of src/egrad_oh3.f (bohr; atoms O,
# H1, H2, H3), planar, V = 6.0909 kcal/mol, one imaginary mode, |dV/dq| ~ 1e-16
# (tests/golden/make_oh3_saddle.py).
OH3_TS = np.array([
    [-1.3546985242473657, -0.2317507462336809, 0.0],
    [-2.098485218890634, 1.4691079024461797, 0.0],
    [0.9504565430735537, -0.3544506279978717, 0.0],
    [2.502727200064447, -0.8829065282146269, 0.0],
])
MASSES_OH3_AMU = [15.999, 1.008, 1.008, 1.008]

# First-order saddle of H + HBr -> H2 + Br on the BrH2 DIM-3C surface of
# src/egrad_brh2.f (bohr; atoms Ha, Br, Hc), collinear Hc-Ha-Br with
# R(Br-Ha) = 2.7216, R(Ha-Hc) = 2.6406, V = 20.996 kcal/mol above Br + H2
# (1.82 above H + HBr), one imaginary mode, |dV/dq| ~ 6e-17
# (tests/golden/make_brh2_saddle.py).
BRH2_TS = np.array([
    [0.0, 0.0, -0.027009331805548],
    [0.0, 0.0, 2.6945795463155617],
    [0.0, 0.0, -2.6675702145100137],
])
MASSES_BRH2_AMU = [1.008, 79.904, 1.008]

# C2v saddle of the O + O2 exchange on the O3 surface of src/egrad_o3.f (bohr;
# atoms O1 (apex), O2, O3): bonds 2.5645 bohr at 105.25 degrees, V = 10.165
# kcal/mol above O + O2(re), one imaginary mode (the antisymmetric stretch),
# |dV/dq| ~ 6e-15 (tests/golden/make_o3_saddle.py).
O3_TS = np.array([
    [-0.6300267594282657, -0.8247489761317799, 0.0],
    [1.9344942909242242, -0.8247489761317799, 0.0],
    [-1.3044675314959586, 1.6494979522635598, 0.0],
])
MASSES_O3_AMU = [15.999, 15.999, 15.999]


def h3_reaction():
    """H_A H_B + H_C -> H_A + H_B H_C (examples/calc_rate/h+h2/rate.key)."""
    return Reaction(bond_form=[(2, 3)], bond_break=[(1, 2)], form_ref=[1.757], break_ref=[1.757],
                    reactants=[[1, 2], [3]], r_inf=16.0, xi_ideal=1.0)


def ch4h_reaction():
    """CH4 + H -> CH3 + H2: C-h1 breaks, h1-Hb forms."""
    rc = CH4H_TS
    return Reaction(bond_form=[(3, 6)], bond_break=[(2, 3)],
                    form_ref=[float(np.linalg.norm(rc[2] - rc[5]))],
                    break_ref=[float(np.linalg.norm(rc[1] - rc[2]))],
                    reactants=[[1, 2, 3, 4, 5], [6]], r_inf=16.0, xi_ideal=1.0)


def oh3_reaction():
    """OH + H2 -> H2O + H: O-H2 forms, H2-H3 breaks, at the saddle's lengths."""
    t = OH3_TS
    return Reaction(bond_form=[(1, 3)], bond_break=[(3, 4)],
                    form_ref=[float(np.linalg.norm(t[0] - t[2]))], break_ref=[float(np.linalg.norm(t[2] - t[3]))],
                    reactants=[[1, 2], [3, 4]], r_inf=16.0, xi_ideal=1.0)


def brh2_reaction():
    """H + HBr -> H2 + Br: Ha-Hc forms, Ha-Br breaks, at the saddle's lengths."""
    t = BRH2_TS
    return Reaction(bond_form=[(1, 3)], bond_break=[(1, 2)],
                    form_ref=[float(np.linalg.norm(t[0] - t[2]))], break_ref=[float(np.linalg.norm(t[0] - t[1]))],
                    reactants=[[1, 2], [3]], r_inf=16.0, xi_ideal=1.0)


def o3_reaction():
    """O3 + O1O2 -> O3O1 + O2: O1-O3 forms, O1-O2 breaks, at the saddle's (equal) lengths."""
    t = O3_TS
    return Reaction(bond_form=[(1, 3)], bond_break=[(1, 2)],
                    form_ref=[float(np.linalg.norm(t[0] - t[2]))], break_ref=[float(np.linalg.norm(t[0] - t[1]))],
                    reactants=[[1, 2], [3]], r_inf=16.0, xi_ideal=1.0)


def mb_momenta(rng, masses, nbeads, beta, shape):
    """Maxwell-Boltzmann bead momenta: p ~ N(0, m nbeads / beta) (andersen.f90:50-56)."""
    sd = np.sqrt(np.asarray(masses)[:, None] * nbeads / np.asarray(beta))  # (natoms, 1)
    return rng.normal(0.0, 1.0, shape) * sd[None, :, :]


def h3_nve(nbeads=16, seed=0, kelvin=300.0, dt_fs=0.1):
    rng = np.random.default_rng(seed)
    masses = mass_au([1.008] * 3)
    beta = float(beta_of(kelvin))
    q = H3_TS[None] + rng.normal(0.0, 0.01, (nbeads, 3, 3))
    p = mb_momenta(rng, masses, nbeads, beta, (nbeads, 3, 3))
    return dict(q=q, p=p, beta=beta, masses=masses, dt=dt_fs * FS_TO_AU, nbeads=nbeads)


def oh3_nve(nbeads=16, seed=0, kelvin=300.0, dt_fs=0.1):
    rng = np.random.default_rng(seed)
    masses = mass_au(MASSES_OH3_AMU)
    beta = float(beta_of(kelvin))
    q = OH3_TS[None] + rng.normal(0.0, 0.01, (nbeads, 4, 3))
    p = mb_momenta(rng, masses, nbeads, beta, (nbeads, 4, 3))
    return dict(q=q, p=p, beta=beta, masses=masses, dt=dt_fs * FS_TO_AU, nbeads=nbeads)


def brh2_nve(nbeads=16, seed=0, kelvin=300.0, dt_fs=0.1):
    rng = np.random.default_rng(seed)
    masses = mass_au(MASSES_BRH2_AMU)
    beta = float(beta_of(kelvin))
    q = BRH2_TS[None] + rng.normal(0.0, 0.01, (nbeads, 3, 3))
    p = mb_momenta(rng, masses, nbeads, beta, (nbeads, 3, 3))
    return dict(q=q, p=p, beta=beta, masses=masses, dt=dt_fs * FS_TO_AU, nbeads=nbeads)


def o3_nve(nbeads=16, seed=0, kelvin=300.0, dt_fs=0.1):
    rng = np.random.default_rng(seed)
    masses = mass_au(MASSES_O3_AMU)
    beta = float(beta_of(kelvin))
    q = O3_TS[None] + rng.normal(0.0, 0.01, (nbeads, 3, 3))
    p = mb_momenta(rng, masses, nbeads, beta, (nbeads, 3, 3))
    return dict(q=q, p=p, beta=beta, masses=masses, dt=dt_fs * FS_TO_AU, nbeads=nbeads)


def ch4h_children(ntraj=4096, nbeads=32, seed=0, kelvin=500.0, dt_fs=0.1):
    rng = np.random.default_rng(seed)
    masses = mass_au([MASS_H_REF, MASS_C_REF, MASS_H_REF, MASS_H_REF, MASS_H_REF, MASS_H_REF])
    beta = float(beta_of(kelvin))
    pert = rng.normal(0.0, 0.02, (ntraj, 1, 6, 3))
    q = np.broadcast_to(CH4H_TS[None, None] + pert, (ntraj, nbeads, 6, 3)).copy()
    q += rng.normal(0.0, 0.02, (ntraj, nbeads, 6, 3)) * 0.25
    p = np.zeros_like(q)
    return dict(q=q, p=p, beta=np.full(ntraj, beta), masses=masses, dt=dt_fs * FS_TO_AU, nbeads=nbeads,
                reaction=ch4h_reaction())


def ragged(ntraj=8192, seed=1, dt_fs=0.1):
    """Bucket list [(surface, nbeads, dict(q, p, beta, masses))] of the ragged batch."""
    from ._abi import CAM_SURFACE_CH4H, CAM_SURFACE_H3
    rng = np.random.default_rng(seed)
    nb_choice = np.array([4, 8, 16, 32, 64])
    surf = np.where(np.arange(ntraj) < ntraj // 2, CAM_SURFACE_H3, CAM_SURFACE_CH4H)
    nbs = nb_choice[rng.integers(0, 5, ntraj)]
    temps = rng.uniform(200.0, 1000.0, ntraj)
    buckets = []
    for s in (CAM_SURFACE_H3, CAM_SURFACE_CH4H):
        for nb in nb_choice:
            sel = np.nonzero((surf == s) & (nbs == nb))[0]
            if len(sel) == 0:
                continue
            n = len(sel)
            beta = beta_of(temps[sel])
            if s == CAM_SURFACE_H3:
                masses = mass_au([1.008] * 3)
                base, na = H3_TS, 3
            else:
                masses = mass_au([MASS_H_REF, MASS_C_REF] + [MASS_H_REF] * 4)
                base, na = CH4H_TS, 6
            q = base[None, None] + rng.normal(0.0, 0.02, (n, int(nb), na, 3))
            sd = np.sqrt(masses[None, None, :, None] * nb / beta[:, None, None, None])
            p = rng.normal(0.0, 1.0, (n, int(nb), na, 3)) * sd
            buckets.append((int(s), int(nb), dict(q=q, p=p, beta=beta, masses=masses, dt=dt_fs * FS_TO_AU,
                                                   traj_index=sel, temps=temps[sel])))
    return buckets
