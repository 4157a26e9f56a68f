"""Write the golden fixtures under tests/golden/ from the reference itself.

Run in the container that holds /root/reference, after `make -C oracle ref`:
    python tests/golden/make_golden.py

* faddeeva_w_table.json: the 57 (z, w(z)) pairs of the reference's Faddeeva
  self-test (fake_spectra/Faddeeva.cpp:1919-2106), numbers only.
* ref_c0.npz: the reference C++ (oracle/_ref/libfsref.so) on the 32^3 / 64
  sightline workload: tau (H I 1215) and column density of 16 sightlines,
  their neighbour lists, and the _near_lines result, with input checksums.
* ref_cases.npz: small hand-built inputs (periodic wrap at every box face,
  duplicate sightlines, all three axes, supports wider than half the box,
  zero and tiny smoothing lengths, cold gas, Gamma = 0, odd and tiny nbins,
  tophat, cubic and quintic kernels) with the reference's outputs.
"""
import json
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from fake_spectra_amd import synth                  # noqa: E402
from fake_spectra_amd.lines import HI_1215, CIV_1548  # noqa: E402
import oracle as O                                   # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
FADDEEVA = "/root/reference/fake_spectra/Faddeeva.cpp"


def faddeeva_table():
    src = open(FADDEEVA).read()
    block = src[src.index("#define NTST 57"):src.index("#undef NTST")]

    def grab(name):
        m = re.search(r"cmplx %s\[NTST\]\s*=\s*\{(.*?)\};" % name, block, re.S)
        body = m.group(1)
        vals = re.findall(r"C\(\s*([^,]+?)\s*,\s*([^)]+?)\s*\)", body)
        return [(float(a), float(b)) for a, b in vals]

    z = grab("z")
    w = grab("w")
    assert len(z) == 57 and len(w) == 57, (len(z), len(w))
    rows = [[a[0], a[1], b[0], b[1]] for a, b in zip(z, w)]
    with open(os.path.join(OUT, "faddeeva_w_table.json"), "w") as f:
        json.dump({"source": "fake_spectra/Faddeeva.cpp:1919-2106 (w(z) self-test, 57 points)",
                   "columns": ["re z", "im z", "re w", "im w"], "rows": rows}, f, indent=0)


def checksums(w):
    return np.array([w.pos.astype(np.float64).sum(), w.hh.astype(np.float64).sum(),
                     w.temp.astype(np.float64).sum(), w.vel.astype(np.float64).sum(),
                     w.dens["H1"].astype(np.float64).sum(), w.cofm.sum()])


def ref_c0(ref):
    w = synth.config(0)
    sel = np.arange(0, 64, 4)
    axis, cofm = w.axis[sel], w.cofm[sel]
    args = list(w.interp_args(HI_1215, True))
    args[16], args[17] = axis, cofm
    tau = ref.interpolate(*args)
    cargs = list(w.interp_args(None, False))
    cargs[16], cargs[17] = axis, cofm
    col = ref.interpolate(*cargs)
    off, parts, dr2 = ref.near_particles(w.box, w.pos, w.hh, axis, cofm)
    near = ref.near_lines(w.box, w.pos, w.hh, w.axis, w.cofm)
    np.savez_compressed(os.path.join(OUT, "ref_c0.npz"), sel=sel, tau=tau, colden=col, offsets=off,
                        parts=parts, dr2=dr2, near_lines=near, checksums=checksums(w))


def crafted_cases():
    """Small inputs built to hit the reference's edge behaviour."""
    rng = np.random.default_rng(20240611)
    box = 1000.0
    cases = {}
    # 1. particles at and across every face of the box, all three axes
    n = 400
    pos = rng.random((n, 3)) * box
    face = rng.integers(0, 3, n)
    side = rng.integers(0, 2, n)
    pos[np.arange(n), face] = np.where(side == 0, rng.random(n) * 3.0, box - rng.random(n) * 3.0)
    pos = pos.astype(np.float32)
    pos[:4] = np.array([[0, 0, 0], [box, box, box], [0.0, box / 2, box], [box / 2, 0, 0]], np.float32)
    pos = np.minimum(pos, np.float32(box))
    hh = (rng.random(n) * 40 + 1).astype(np.float32)
    hh[5] = 600.0        # wider than half the box
    hh[6] = 0.0          # zero support
    hh[7] = 1e-3
    ng = 4
    g = (np.arange(ng) + 0.5) * box / ng
    gg = np.array([[x, y] for x in g for y in g])
    axis = np.concatenate([np.full(len(gg), a) for a in (1, 2, 3)]).astype(np.int32)
    cofm = np.zeros((3 * len(gg), 3))
    cofm[: len(gg), 1:] = gg
    cofm[len(gg): 2 * len(gg)][:, [0, 2]] = gg
    cofm[2 * len(gg):, :2] = gg
    # edge sightlines: on the faces, duplicated, and at exactly 0 / box
    extra_axis = np.array([1, 1, 2, 3, 3, 1], np.int32)
    extra = np.array([[0, 0, 0], [0, 0, 0], [box, 0, box], [0, box, 0], [1.5, 998.5, 0], [0, 500, 999.9]])
    axis = np.concatenate([axis, extra_axis])
    cofm = np.concatenate([cofm, extra])
    vel = rng.normal(0, 30, (n, 3)).astype(np.float32)
    temp = (10 ** rng.uniform(0.5, 6.5, n)).astype(np.float32)   # includes very cold gas
    dens = (10 ** rng.uniform(-8, -3, n) * 3.0857e21).astype(np.float32)
    cases["wrap"] = dict(box=box, pos=pos, hh=hh, axis=axis, cofm=cofm, vel=vel, temp=temp, dens=dens)
    # 2. a dense clump on one sightline with strong absorption (long damping wings)
    n2 = 150
    pos2 = (np.array([500.0, 500.0, 500.0]) + rng.normal(0, 20, (n2, 3))).astype(np.float32)
    pos2[:, 0] = (rng.random(n2) * box).astype(np.float32)
    hh2 = (rng.random(n2) * 30 + 10).astype(np.float32)
    vel2 = rng.normal(0, 50, (n2, 3)).astype(np.float32)
    temp2 = (10 ** rng.uniform(2, 5, n2)).astype(np.float32)
    dens2 = (10 ** rng.uniform(-4, -1, n2) * 3.0857e21).astype(np.float32)
    axis2 = np.array([1, 1, 1], np.int32)
    cofm2 = np.array([[0, 500, 500], [0, 505, 498], [0, 540, 470]], np.float64)
    cases["clump"] = dict(box=box, pos=pos2, hh=hh2, axis=axis2, cofm=cofm2, vel=vel2, temp=temp2, dens=dens2)
    return cases


def ref_cases(ref):
    velfac = 0.1115235417541224
    out = {}
    for name, c in crafted_cases().items():
        for kern in (0, 1, 3):
            for (ln, gam_scale, nb) in ((HI_1215, 1.0, 257), (HI_1215, 0.0, 64), (CIV_1548, 1.0, 300), (HI_1215, 1.0, 7)):
                key = "%s_k%d_%s_g%g_n%d" % (name, kern, ln.name.replace(" ", ""), gam_scale, nb)
                tau = ref.interpolate(1, nb, kern, c["box"], velfac, 0.25, ln.lambda_cm, ln.gamma * gam_scale,
                                      ln.fosc, ln.amumass, 1e-7, c["pos"], c["vel"], c["dens"], c["temp"], c["hh"],
                                      c["axis"], c["cofm"])
                out[key + "_tau"] = tau
            col = ref.interpolate(0, 257, kern, c["box"], velfac, 0.25, HI_1215.lambda_cm, HI_1215.gamma,
                                  HI_1215.fosc, HI_1215.amumass, 1e-7, c["pos"], np.zeros(1, np.float32), c["dens"],
                                  np.zeros(1, np.float32), c["hh"], c["axis"], c["cofm"])
            out["%s_k%d_colden" % (name, kern)] = col
        off, parts, dr2 = ref.near_particles(c["box"], c["pos"], c["hh"], c["axis"], c["cofm"])
        out[name + "_offsets"], out[name + "_parts"], out[name + "_dr2"] = off, parts, dr2
        out[name + "_near"] = ref.near_lines(c["box"], c["pos"], c["hh"], c["axis"], c["cofm"])
        for k, v in c.items():
            out["%s_in_%s" % (name, k)] = np.asarray(v)
    out["velfac"] = np.array(velfac)
    np.savez_compressed(os.path.join(OUT, "ref_cases.npz"), **out)


if __name__ == "__main__":
    faddeeva_table()
    ref = O.load_ref()
    ref_c0(ref)
    ref_cases(ref)
    for f in sorted(os.listdir(OUT)):
        print(f, os.path.getsize(os.path.join(OUT, f)))
