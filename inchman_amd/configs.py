"""Synthetic workloads of the benchmark and the parity tests (seeded, no data files).

  plain_gmp      32x32 homogeneous GMP, 3 species A,B,C (0->A, 0->B, A+B->C, C->0)
  schnakenberg   256x256 stochastic inhomogeneous drift-diffusion, Schnakenberg U,V
  ensemble       512 realizations of plain_gmp's network on square grids with sides
                 log-uniform in [64,512], rounded to multiples of 16
  particles      10^7 particles on a 1024x1024 periodic domain
  rdd            1024x1024 Schnakenberg reaction-diffusion in concentration form
"""
import math

import numpy as np

from .gpgmp import DiffusionModel, STOCHASTIC_HOMOGENEOUS, STOCHASTIC_INHOMOGENEOUS


# ---- plain GMP (32x32, h = 1 um, reflective) ----------------------------------------
def plain_gmp_model(width=32, height=None, h=1.0):
    height = width if height is None else height
    # dt = 0.1 h^2/(4 D_max) = 0.025 s equals the reference's p0/6 h^2/D_max at p0 = 0.15
    m = DiffusionModel(width, height, physical_length=h * width, p0=0.15,
                       solver=STOCHASTIC_HOMOGENEOUS)
    a = m.addSpecies("A", 1.0)
    b = m.addSpecies("B", 0.5)
    c = m.addSpecies("C", 0.1)
    m.addReaction(1.0, products={a: 1}, id="A_in")
    m.addReaction(1.0, products={b: 1}, id="B_in")
    m.addReaction(0.01, reactants={a: 1, b: 1}, products={c: 1}, id="AB")
    m.addReaction(0.1, reactants={c: 1}, id="C_out")
    return m


def plain_gmp_state(width=32, height=None, seed=12345):
    height = width if height is None else height
    rng = np.random.default_rng(seed)
    st = np.zeros((3, height, width), np.int32)
    st[0] = rng.poisson(50, (height, width))
    st[1] = rng.poisson(50, (height, width))
    return st


# ---- Schnakenberg inhomogeneous drift-diffusion (256x256) --------------------------
SCHNAKENBERG_K = (1.0, 3.0, 1e-4, 1.0)


def schnakenberg_fields(width, height, h=1.0, D0=(0.05, 1.0), drift=(0.02, 0.0)):
    """D_s(x,y) = D0_s (1 + 0.5 sin(2 pi x / L)) with the reference's PhysicalX
    convention x = L/W (i - W/2) (cl_include/helpers.cl:29-32); constant drift."""
    L = h * width
    i = np.arange(width, dtype=np.float64)
    x = (np.float32(L) / np.float32(width)) * (i.astype(np.float32) - np.float32(width) / np.float32(2.0))
    prof = 1.0 + 0.5 * np.sin(2.0 * math.pi * x.astype(np.float64) / L)
    S = len(D0)
    D = np.empty((S, height, width), np.float32)
    for s in range(S):
        D[s] = (D0[s] * prof)[None, :].astype(np.float32)
    rx = np.full((S, height, width), drift[0], np.float32)
    ry = np.full((S, height, width), drift[1], np.float32)
    return D, D.copy(), rx, ry


def schnakenberg_model(width=256, height=None, h=1.0):
    height = width if height is None else height
    m = DiffusionModel(width, height, physical_length=h * width, p0=0.15,
                       solver=STOCHASTIC_INHOMOGENEOUS)
    u = m.addSpecies("U", 0.0)
    v = m.addSpecies("V", 0.0)
    k1, k2, k3, k4 = SCHNAKENBERG_K
    m.addReaction(k1, products={u: 1}, id="U_in")
    m.addReaction(k2, products={v: 1}, id="V_in")
    m.addReaction(k3, reactants={u: 2, v: 1}, products={u: 3}, id="autocatalysis")
    m.addReaction(k4, reactants={u: 1}, id="U_out")
    m.setDriftDiffusivityFields(*schnakenberg_fields(width, height, h))
    return m


def schnakenberg_state(width=256, height=None, seed=12345):
    height = width if height is None else height
    rng = np.random.default_rng(seed)
    st = np.zeros((2, height, width), np.int32)
    st[0] = rng.poisson(200, (height, width))
    st[1] = rng.poisson(75, (height, width))
    return st


# ---- ensemble ---------------------------------------------------------------------------
def ensemble_sides(n=512, lo=64, hi=512, seed=2024):
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(math.log(lo), math.log(hi), n))
    return (16 * np.maximum(1, np.round(s / 16))).astype(np.int32)


def threefry2x32_20(c0, c1, k0, k1):
    """Threefry-2x32 with 20 rounds (Salmon et al., SC'11: rotations 13 15 26 6 17 29 16 24,
    key parity 0x1BD11BDA, a key injection after every fourth round), vectorised over numpy
    uint32 arrays; the generator of the solver kernels, used here to derive seeds."""
    M = np.uint64(0xFFFFFFFF)
    c0, c1, k0, k1 = (np.asarray(v, np.uint64) & M for v in (c0, c1, k0, k1))
    ks = (k0, k1, (k0 ^ k1 ^ np.uint64(0x1BD11BDA)) & M)
    rot = (13, 15, 26, 6, 17, 29, 16, 24)
    x0, x1 = (c0 + ks[0]) & M, (c1 + ks[1]) & M
    for r in range(20):
        x0 = (x0 + x1) & M
        s = rot[r % 8]
        x1 = ((x1 << np.uint64(s)) | (x1 >> np.uint64(32 - s))) & M
        x1 ^= x0
        if r % 4 == 3:
            j = r // 4 + 1
            x0 = (x0 + ks[j % 3]) & M
            x1 = (x1 + ks[(j + 1) % 3] + np.uint64(j)) & M
    return x0.astype(np.uint32), x1.astype(np.uint32)


def ensemble_seeds(n=512, key=0xBA7C):
    """Realization seeds s_r = word 0 of Threefry-2x32-20 at counter (r, 0), key (0xBA7C, 0)."""
    r = np.arange(n, dtype=np.uint64)
    return threefry2x32_20(r, 0, key, 0)[0]


def ensemble_state(side, r, seed=12345):
    return plain_gmp_state(int(side), seed=seed + 7919 * r)


# ---- differentiable reaction-diffusion (1024x1024, 2000 steps) -------------------------
RDD_D0 = (0.05, 1.0)


def rdd_problem(width=1024, height=None, steps=2000, seed=7, target_seed=11, field="smoothed_lognormal"):
    """Schnakenberg (k1..k4 of the stochastic configuration) in concentration form on a
    seeded diffusivity field.  field "smoothed_lognormal" (the configured one):
    D = G_8 * exp(N(0, 0.3^2)), i.e. per-cell log-normal noise smoothed by a Gaussian of
    sigma 8 cells; "lognormal_grf": D = exp(0.3 g) with g a unit-variance Gaussian random
    field of correlation length 8 cells (white noise smoothed, then rescaled), a rougher
    field.  dt = 0.2 h^2 / (4 max_s D0_s max D).  Initial u, v near 200 / 75; the target is
    a seeded field of the magnitude the run reaches."""
    from scipy.ndimage import gaussian_filter

    height = width if height is None else height
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((height, width))
    if field == "smoothed_lognormal":
        D = gaussian_filter(np.exp(0.3 * noise), 8.0, mode="reflect").astype(np.float32)
    elif field == "lognormal_grf":
        g = gaussian_filter(noise, 8.0, mode="reflect")
        g = (g - g.mean()) / g.std()
        D = np.exp(0.3 * g).astype(np.float32)
    else:
        raise ValueError("unknown field %r" % (field,))
    dt = float(np.float32(0.2 / (4.0 * max(RDD_D0) * float(D.max()))))
    u0 = (200.0 + np.sqrt(200.0) * rng.standard_normal((height, width))).astype(np.float32)
    v0 = (75.0 + np.sqrt(75.0) * rng.standard_normal((height, width))).astype(np.float32)
    trng = np.random.default_rng(target_seed)
    ut = (1.0 + 0.1 * trng.standard_normal((height, width))).astype(np.float32)
    vt = (60.0 + 5.0 * trng.standard_normal((height, width))).astype(np.float32)
    return {"D": D, "u0": u0, "v0": v0, "ut": ut, "vt": vt, "dt": dt, "steps": steps,
            "k": SCHNAKENBERG_K, "D0": RDD_D0, "h": 1.0, "field": field}
