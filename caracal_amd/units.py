"""Unit conventions of the reference (src/general.f90:250-264, src/dynamic.f90:79,584,589)."""
import numpy as np

EMASS_AMU = 5.485799095e-4                 # electron mass in amu (general.f90)
AMU_TO_ME = 1.0 / EMASS_AMU
KB_HARTREE = 0.316679e-5                   # Boltzmann constant in Eh/K (dynamic.f90:79)
# dt = dt_fs / 2.41888428E-2: the divisor is a default-REAL literal, so it is
# rounded to binary32 before the division (dynamic.f90:589)
FS_TO_AU = 1.0 / float(np.float32(2.41888428e-2))
BOHR_ANGSTROM = 0.52917721092

MASS_H_REF = 1.00782503207                 # atommass.f90:59
MASS_C_REF = 12.0


def mass_au(amu):
    return np.asarray(amu, dtype=np.float64) / EMASS_AMU


def beta_of(kelvin):
    return 1.0 / (np.asarray(kelvin, dtype=np.float64) * KB_HARTREE)
