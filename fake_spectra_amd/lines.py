"""Absorption-line constants and the reference's physical constants.

Line data are VPFIT atom.dat values as the reference reads them
(fake_spectra/line_data.py:59-111): rest wavelength [Angstrom], oscillator
strength, damping constant Gamma [1/s].  Species masses follow
fake_spectra/line_data.py:21.
"""
from dataclasses import dataclass
import math

# fake_spectra/absorption.cpp:21-33, absorption.h:4
SIGMA_T = 6.652458558e-25      # Thomson cross-section [cm^2]
BOLTZMANN = 1.3806504e-16      # [erg/K]
LIGHT = 2.99792458e10          # [cm/s]
PROTONMASS = 1.67262178e-24    # [g]

SPECIES_MASS = {"H": 1.00794, "He": 4.002602, "C": 12.011, "N": 14.00674, "O": 15.9994,
                "Ne": 20.18, "Mg": 24.3050, "Si": 28.0855, "Fe": 55.847}


@dataclass(frozen=True)
class Line:
    """One transition: the (lambda, fosc, gamma, amumass) quadruple that
    _Particle_Interpolate takes (spectra.py:666-673)."""
    name: str
    elem: str
    ion: int
    lambda_A: float
    fosc: float
    gamma: float

    @property
    def amumass(self) -> float:
        return SPECIES_MASS[self.elem]

    @property
    def lambda_cm(self) -> float:
        # spectra.py:673 passes line.lambda_X*1e-8
        return self.lambda_A * 1e-8


# atom.dat:7, atom.dat:8, atom.dat:174
HI_1215 = Line("H I 1215", "H", 1, 1215.6701, 0.416400, 6.265e8)
HI_1025 = Line("H I 1025", "H", 1, 1025.7223, 0.079120, 1.897e8)
CIV_1548 = Line("C IV 1548", "C", 4, 1548.2049, 0.189900, 2.642e8)

LINES = {"HI1215": HI_1215, "HI1025": HI_1025, "CIV1548": CIV_1548}


def line_factors(lambda_cm: float, gamma: float, fosc: float, amumass: float):
    """Per-line constants of LineAbsorption (absorption.cpp:152-161):
    sigma_a [cm^2], bfac [km/s/sqrt(K)], voigt_fac [km/s]."""
    sigma_a = math.sqrt(3.0 * math.pi * SIGMA_T / 8.0) * lambda_cm * fosc
    bfac = math.sqrt(2.0 * BOLTZMANN / (amumass * PROTONMASS)) / 1e5
    voigt_fac = gamma * lambda_cm / (4.0 * math.pi) / 1e5
    return sigma_a, bfac, voigt_fac
