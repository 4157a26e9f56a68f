"""Seeded synthetic gas fields and sightline tables in the reference's units.

The reference reads particles from snapshots and converts them to the
arguments of _Particle_Interpolate in Spectra._read_particle_data
(spectra.py:550-617) and Spectra.__init__ (spectra.py:210-241).  No snapshot
exists here, so this module draws the gas fields from seeded numpy generators
and then applies the same unit pipeline:

* rscale = UnitLength_in_cm * a / h, rounded to float32 (spectra.py:210)
* H(z) = 100 h sqrt(Om/a^3 + OL); velfac = rscale H(z) / 3.085678e24 (spectra.py:213-216)
* vmax = box * velfac; dvbin = vmax / nbins (spectra.py:217-232)
* n_gas [amu/cm^3] = rho * UnitDensity_in_cgs h^2 (1+z)^3 / m_p, with the
  conversion factor rounded to float32 (gas_properties.py:105-110)
* elem_den = n_gas * rscale * mass_frac [* x_ion] / amumass (spectra.py:593-615)
* temperatures <= 0 are clamped to 1 K (spectra.py:588-589)

Physical model of the synthetic field (an input choice, not reference
behaviour): lognormal overdensity, SPH smoothing length from the particle mass
and density with 32 neighbours, power-law temperature-density relation,
photoionisation-equilibrium neutral fraction, Gaussian peculiar velocities.
The clustered variant adds periodic Gaussian blobs whose particles take the
blob's own density profile as an extra overdensity factor.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math
from typing import Dict, List, Optional

import numpy as np

from .lines import Line, HI_1215, HI_1025, CIV_1548, SPECIES_MASS, PROTONMASS

UNIT_LENGTH_CM = 3.085678e21       # 1 kpc/h  (unitsystem.py:7)
UNIT_MASS_G = 1.98892e43           # 1e10 Msun/h
RHO_CRIT_INTERNAL = 27.755e-9      # rho_crit in 1e10 Msun/h / (kpc/h)^3 (h-scaled)
HYDROGEN_MASS_FRAC = 0.76          # spectra.py:702 primordial fallback
SOLAR_C_BY_NUMBER = 2.69e-4        # spectra.py:245


@dataclass
class Cosmology:
    redshift: float = 3.0
    hubble: float = 0.7
    omega_m: float = 0.3
    omega_l: float = 0.7
    omega_b: float = 0.045

    @property
    def atime(self) -> float:
        return 1.0 / (1.0 + self.redshift)

    def rscale(self) -> float:
        return float(np.float32(UNIT_LENGTH_CM * self.atime / self.hubble))

    def hz(self) -> float:
        return 100.0 * self.hubble * math.sqrt(self.omega_m / self.atime ** 3 + self.omega_l)

    def velfac(self) -> float:
        return self.rscale() * self.hz() / 3.085678e24

    def density_conv(self) -> float:
        """Internal density -> physical amu/cm^3 (gas_properties.py:105-110)."""
        unit_density = UNIT_MASS_G / UNIT_LENGTH_CM ** 3
        return float(np.float32(unit_density * self.hubble ** 2 / PROTONMASS * (1 + self.redshift) ** 3))


@dataclass
class Workload:
    """Everything _Particle_Interpolate needs, for one or more lines."""
    name: str
    box: float
    nbins: int
    kernel: int
    cosmo: Cosmology
    pos: np.ndarray            # float32 [N,3], kpc/h comoving
    vel: np.ndarray            # float32 [N,3], km/s physical
    temp: np.ndarray           # float32 [N], K
    hh: np.ndarray             # float32 [N], kernel support radius, kpc/h
    overdensity: np.ndarray    # float32 [N], 1+delta
    neutral_fraction: np.ndarray   # float32 [N], x_HI
    axis: np.ndarray           # int32 [L], 1..3
    cofm: np.ndarray           # float64 [L,3]
    lines: List[Line]
    dens: Dict[str, np.ndarray] = field(default_factory=dict)   # per species key, float32 [N]
    tautail: float = 1e-7      # spectra.py:135

    @property
    def velfac(self) -> float:
        return self.cosmo.velfac()

    @property
    def atime(self) -> float:
        return self.cosmo.atime

    @property
    def npart(self) -> int:
        return int(self.pos.shape[0])

    @property
    def nlos(self) -> int:
        return int(self.axis.shape[0])

    def dens_for(self, line: Line) -> np.ndarray:
        return self.dens[species_key(line)]

    def interp_args(self, line: Optional[Line], compute_tau: bool):
        """Positional argument tuple of _Particle_Interpolate for one line
        (spectra.py:673).  Column density uses H I 1215 constants and the
        species density, as the reference does (spectra.py:541-545)."""
        ln = line if line is not None else HI_1215
        dens = self.dens_for(ln)
        vel = self.vel if compute_tau else np.zeros(1, dtype=np.float32)
        temp = self.temp if compute_tau else np.zeros(1, dtype=np.float32)
        return (int(compute_tau), self.nbins, self.kernel, self.box, self.velfac, self.atime,
                ln.lambda_cm, ln.gamma, ln.fosc, ln.amumass, self.tautail,
                self.pos, vel, dens, temp, self.hh, self.axis, self.cofm)


def species_key(line: Line) -> str:
    return "%s%d" % (line.elem, line.ion)


def _uniform_positions(rng, n, box):
    pos = rng.random((n, 3), dtype=np.float64) * box
    pos = pos.astype(np.float32)
    # float32 rounding can land exactly on the box edge
    np.minimum(pos, np.nextafter(np.float32(box), np.float32(0)), out=pos)
    return pos


def _clustered_positions(rng, n, box, nblobs=2000, blob_frac=0.3, sigma=300.0, slope=-1.8,
                         mmin=1.0, mmax=100.0):
    """70% uniform + 30% in periodic Gaussian blobs with power-law masses.
    Returns (pos float32 [n,3], blob overdensity factor float64 [n])."""
    centres = rng.random((nblobs, 3)) * box
    # inverse-CDF draw from p(M) ~ M^slope on [mmin, mmax]
    e = slope + 1.0
    u = rng.random(nblobs)
    masses = (mmin ** e + u * (mmax ** e - mmin ** e)) ** (1.0 / e)
    nblob_part = int(round(blob_frac * n))
    nfield = n - nblob_part
    counts = rng.multinomial(nblob_part, masses / masses.sum())
    owner = np.repeat(np.arange(nblobs), counts)
    offs = rng.normal(0.0, sigma, (nblob_part, 3))
    bpos = np.mod(centres[owner] + offs, box)
    fpos = rng.random((nfield, 3)) * box
    pos = np.concatenate([fpos, bpos]).astype(np.float32)
    np.minimum(pos, np.nextafter(np.float32(box), np.float32(0)), out=pos)
    # blob density relative to the mean at the particle's radius
    mean_num = n / box ** 3
    r2 = np.sum(offs * offs, axis=1)
    prof = counts[owner] * (2 * math.pi * sigma * sigma) ** -1.5 * np.exp(-0.5 * r2 / (sigma * sigma))
    factor = np.ones(n)
    factor[nfield:] = np.maximum(1.0, prof / mean_num)
    # shuffle so particle index order carries no spatial information
    perm = rng.permutation(n)
    return pos[perm], factor[perm]


def make_gas(name, n_side, box, sigma_ln, nbins, kernel, lines, axis, cofm, seed,
             clustered=False, cosmo=None, metallicity=1e-3, civ_fraction=0.1, npart=None):
    """Draw one seeded gas field and convert it to _Particle_Interpolate units."""
    cosmo = cosmo or Cosmology()
    rng = np.random.Generator(np.random.PCG64(seed))
    n = npart if npart is not None else n_side ** 3
    if clustered:
        pos, blob = _clustered_positions(rng, n, box)
    else:
        pos, blob = _uniform_positions(rng, n, box), None
    g = rng.standard_normal(n)
    one_plus_delta = np.exp(sigma_ln * g - 0.5 * sigma_ln ** 2)
    if blob is not None:
        one_plus_delta *= blob
    rho_bar = cosmo.omega_b * RHO_CRIT_INTERNAL            # internal units
    mass = rho_bar * box ** 3 / n
    rho = rho_bar * one_plus_delta
    hh = (32.0 * mass / (4.0 / 3.0 * math.pi * rho)) ** (1.0 / 3.0)
    temp = 1e4 * one_plus_delta ** 0.6
    x_hi = np.minimum(1.0, 1e-5 * one_plus_delta * (temp / 1e4) ** -0.7)
    vel = rng.normal(0.0, 30.0, (n, 3)).astype(np.float32)
    temp32 = temp.astype(np.float32)
    temp32[temp32 <= 0] = 1.0
    # density pipeline, float32 as in spectra.py:576-615
    den = (rho.astype(np.float32) * np.float32(cosmo.density_conv())).astype(np.float32)
    base = den * np.float32(cosmo.rscale())
    dens = {}
    for ln in lines:
        key = species_key(ln)
        if key in dens:
            continue
        if ln.elem == "H":
            d = base * np.float32(HYDROGEN_MASS_FRAC) * x_hi.astype(np.float32)
        else:
            # metal mass fraction: metallicity x solar abundance by mass, fixed ion fraction
            solar_mass_frac = HYDROGEN_MASS_FRAC * SOLAR_C_BY_NUMBER * SPECIES_MASS["C"] / SPECIES_MASS["H"]
            d = base * np.float32(metallicity * solar_mass_frac) * np.float32(civ_fraction)
        dens[key] = (d / np.float32(ln.amumass)).astype(np.float32)
    w = Workload(name=name, box=float(box), nbins=int(nbins), kernel=int(kernel), cosmo=cosmo,
                 pos=pos, vel=vel, temp=temp32, hh=hh.astype(np.float32),
                 overdensity=one_plus_delta.astype(np.float32),
                 neutral_fraction=x_hi.astype(np.float32),
                 axis=np.ascontiguousarray(axis, dtype=np.int32),
                 cofm=np.ascontiguousarray(cofm, dtype=np.float64), lines=list(lines), dens=dens)
    return w


def random_sightlines(rng, nlos, box, axes=(1,)):
    axis = rng.choice(np.asarray(axes, dtype=np.int32), size=nlos).astype(np.int32)
    cofm = rng.random((nlos, 3)) * box
    return axis, cofm


def gridded_sightlines(nspec, box):
    """Three-axis regular grid (griddedspectra.py:59-88 with axis < 0)."""
    nn, mm = np.meshgrid(np.arange(nspec), np.arange(nspec), indexing="ij")
    nn = nn.ravel()
    mm = mm.ravel()
    z = np.zeros_like(nn)
    grid = np.concatenate([np.stack([z, nn, mm], 1), np.stack([nn, z, mm], 1),
                           np.stack([nn, mm, z], 1)])
    axis = np.concatenate([np.full(nspec * nspec, 1), np.full(nspec * nspec, 2),
                           np.full(nspec * nspec, 3)]).astype(np.int32)
    dx = box / (1.0 * nspec)
    return axis, (dx * grid).astype(np.float64)


# ---------------------------------------------------------------------------
# The five workloads (32^3 ... 512^3).  `scale` shrinks particle and sightline
# counts for quick tests while keeping box, density law and pixel count.

def config(idx: int, scale: float = 1.0, seed: int = 1234, nlos: Optional[int] = None,
           n_side: Optional[int] = None) -> Workload:
    srng = np.random.Generator(np.random.PCG64(seed + 7919 * (idx + 1)))
    if idx == 0:
        box, ns, sig, nb, kern = 20000.0, 32, 1.0, 1024, 1
        L = nlos or max(1, int(round(64 * scale)))
        axis, cofm = random_sightlines(srng, L, box, (1,))
        lines = [HI_1215]
        clustered = False
    elif idx == 1:
        box, ns, sig, nb, kern = 60000.0, 256, 1.5, 2048, 1
        L = nlos or max(1, int(round(5000 * scale)))
        axis, cofm = random_sightlines(srng, L, box, (1,))
        lines = [HI_1215, HI_1025]
        clustered = False
    elif idx == 2:
        box, ns, sig, nb, kern = 60000.0, 256, 1.5, 1024, 1
        nspec = max(2, int(round(224 * math.sqrt(scale)))) if nlos is None else nlos
        axis, cofm = gridded_sightlines(nspec, box)
        lines = [HI_1215, HI_1025, CIV_1548]
        clustered = True
    elif idx == 3:
        box, ns, sig, nb, kern = 60000.0, 128, 1.5, 1024, 1
        L = nlos or max(1, int(round(1024 * scale)))
        axis, cofm = random_sightlines(srng, L, box, (1,))
        lines = [HI_1215]
        clustered = False
    elif idx == 4:
        box, ns, sig, nb, kern = 120000.0, 512, 1.5, 2048, 0
        L = nlos or max(1, int(round(100000 * scale)))
        axis, cofm = random_sightlines(srng, L, box, (1, 2, 3))
        lines = [HI_1215]
        clustered = True
    else:
        raise ValueError("config index must be 0..4")
    if n_side is None:
        n_side = max(4, int(round(ns * scale ** (1.0 / 3.0)))) if scale != 1.0 else ns
    return make_gas("c%d" % idx, n_side, box, sig, nb, kern, lines, axis, cofm,
                    seed=seed + idx, clustered=clustered)
