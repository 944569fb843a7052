"""Synthetic inputs for the benchmark and parity workloads (host numpy).

Random draws use numpy's PCG64 Generator seeded per workload, evaluated at the
node coordinates so that duplicated (coincident) nodes receive identical
values and fields are continuous across element faces.
"""
import numpy as np


def gaussian_bumps_2d(x, y, seed=12345, nbumps=8, ncomp=2, L=1.0):
    """Sum of Gaussian bumps exp(-(r/w)^2), periodic minimum-image distance.

    Centres ~ U[0,L)^2, widths ~ U[0.03,0.1], amplitudes ~ N(0,1), drawn
    independently per component (2-D Burgers workload, 16x16 elements, p=4).
    """
    rng = np.random.default_rng(seed)
    comps = []
    for _ in range(ncomp):
        cx = rng.uniform(0.0, L, nbumps)
        cy = rng.uniform(0.0, L, nbumps)
        w = rng.uniform(0.03, 0.1, nbumps)
        a = rng.standard_normal(nbumps)
        u = np.zeros_like(x)
        for b in range(nbumps):
            dx = x - cx[b]
            dy = y - cy[b]
            dx -= L * np.round(dx / L)
            dy -= L * np.round(dy / L)
            u += a[b] * np.exp(-(dx * dx + dy * dy) / (w[b] * w[b]))
        comps.append(u)
    return comps


def fourier_field_3d(x, y, z, seed=12345, kmin=1, kmax=6, ncomp=3, L=1.0, xp=np):
    """Random periodic Fourier field: integer wavevectors with kmin<=|k|<=kmax,
    amplitude |k|^-2, seeded phases, independent per component.  xp is the
    array module of x, y, z (numpy, or torch for device tensors)."""
    rng = np.random.default_rng(seed)
    ks = []
    for a in range(-kmax, kmax + 1):
        for b in range(-kmax, kmax + 1):
            for c in range(0, kmax + 1):
                kk = np.sqrt(a * a + b * b + c * c)
                if kmin <= kk <= kmax and (c > 0 or b > 0 or (b == 0 and a > 0)):
                    ks.append((a, b, c, kk))
    comps = []
    tp = 2.0 * np.pi / L
    for _ in range(ncomp):
        ph = rng.uniform(0.0, 2.0 * np.pi, len(ks))
        u = xp.zeros_like(x)
        for (a, b, c, kk), f in zip(ks, ph):
            u += kk ** -2 * xp.cos(tp * (a * x + b * y + c * z) + f)
        comps.append(u / np.sqrt(len(ks)))
    return comps


def unique_normal(grid, seed=67890):
    """N(0,1) per unique node, scattered to all duplicates through the GGFX
    groups (each group takes the draw of its smallest member)."""
    rng = np.random.default_rng(seed)
    offs, mem = grid.ggfx_offs, grid.ggfx_members
    n = offs.size - 1
    draw = rng.standard_normal(n)
    first = mem[offs[:-1]]
    return draw[first]
