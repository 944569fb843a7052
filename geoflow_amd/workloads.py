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


def box_first_member(grid):
    """Smallest member of every node's GGFX group on a 3-D box whose maps
    are left to the device (BoxGrid(..., device_maps=True)): in each
    direction where the node has a partner (index 0 or N-1 on a shared or
    periodic face) the group offers two copies, and the global index
    e*N^3 + local is smallest for the copy with the smaller element
    coordinate (the element index dominates), or the smaller node index
    when the partner is in the same element (one periodic element)."""
    N = grid.N[0]
    ne = grid.ne
    NN = N ** 3
    q = np.arange(grid.ndof, dtype=np.int64)
    e, r = q // NN, q % NN
    node = [r % N, (r // N) % N, r // (N * N)]
    ec = [e % ne[0], (e // ne[0]) % ne[1], e // (ne[0] * ne[1])]
    best_c, best_n = [], []
    for d in range(3):
        c, n = ec[d].copy(), node[d].copy()
        per = bool(grid.periodic[d])
        lo = (node[d] == 0) & ((ec[d] > 0) | per)
        hi = (node[d] == N - 1) & ((ec[d] < ne[d] - 1) | per)
        pc = np.where(lo, np.where(ec[d] > 0, ec[d] - 1, ne[d] - 1),
                      np.where(hi, np.where(ec[d] < ne[d] - 1, ec[d] + 1, 0), ec[d]))
        pn = np.where(lo, N - 1, np.where(hi, 0, node[d]))
        take = (pc < c) | ((pc == c) & (pn < n))
        best_c.append(np.where(take, pc, c))
        best_n.append(np.where(take, pn, n))
    eb = best_c[0] + ne[0] * (best_c[1] + ne[1] * best_c[2])
    return eb * NN + best_n[0] + N * (best_n[1] + N * best_n[2])


def unique_normal(grid, seed=67890):
    """N(0,1) per unique node, scattered to all duplicates through the GGFX
    groups (each group takes the draw of its smallest member)."""
    rng = np.random.default_rng(seed)
    if getattr(grid, "ggfx_offs", None) is None and getattr(grid, "device_maps", False):
        first = box_first_member(grid)
        draw = rng.standard_normal(grid.ndof)
        return draw[first]
    offs, mem = grid.ggfx_offs, grid.ggfx_members
    n = offs.size - 1
    draw = rng.standard_normal(n)
    first = mem[offs[:-1]]
    return draw[first]


def williamson1_velocity(x, y, z, c0, alpha):
    """Solid-body rotation of Williamson et al. (JCP 102, 1992) test 1:
    latitude/longitude components
        v_lat = -c0 sin(lon) sin(alpha)
        v_lon =  c0 (cos(lat) cos(alpha) + sin(lat) cos(lon) sin(alpha))
    turned into Cartesian components as GMTK::vsphere2cart does for physical
    components (gmtk.ipp:3019-3032); this is rotation about the axis
    (-sin alpha, 0, cos alpha) at angular speed c0 / R.  The reference's
    set-up (ginitstate_direct_user.ipp:582-593) has the opposite sign on the
    sin(lat) cos(lon) sin(alpha) term, which is not a rigid rotation for
    alpha != 0 (the bell would not return after a revolution); the published
    form is used so that the exact solution is the rotated bell."""
    r = np.sqrt(x * x + y * y + z * z)
    lat = np.arcsin(z / r)
    lon = np.arctan2(y, x)
    vlat = -c0 * np.sin(lon) * np.sin(alpha)
    vlon = c0 * (np.cos(lat) * np.cos(alpha) + np.sin(lat) * np.cos(lon) * np.sin(alpha))
    vx = -vlat * np.sin(lat) * np.cos(lon) - vlon * np.sin(lon)
    vy = -vlat * np.sin(lat) * np.sin(lon) + vlon * np.cos(lon)
    vz = vlat * np.cos(lat)
    return [vx, vy, vz]


def cosine_bell(x, y, z, lat0, lon0, radius=1.0, h0=1.0):
    """Williamson test-1 cosine bell of radius R/3 centred at (lat0, lon0):
    h = h0/2 (1 + cos(pi r / (R/3))) for great-circle distance r < R/3."""
    R = radius
    rr = np.sqrt(x * x + y * y + z * z)
    lat = np.arcsin(np.clip(z / rr, -1.0, 1.0))
    lon = np.arctan2(y, x)
    c = np.sin(lat0) * np.sin(lat) + np.cos(lat0) * np.cos(lat) * np.cos(lon - lon0)
    r = R * np.arccos(np.clip(c, -1.0, 1.0))
    rb = R / 3.0
    return np.where(r < rb, 0.5 * h0 * (1.0 + np.cos(np.pi * r / rb)), 0.0)


def rotated_bell(x, y, z, lat0, lon0, alpha, angle, radius=1.0, h0=1.0):
    """Exact solution of the rotation: the initial bell evaluated at the
    points rotated back by `angle` about the axis (-sin alpha, 0, cos alpha)
    (Rodrigues' formula)."""
    n = np.array([-np.sin(alpha), 0.0, np.cos(alpha)])
    c, s = np.cos(-angle), np.sin(-angle)
    dot = n[0] * x + n[1] * y + n[2] * z
    cx = n[1] * z - n[2] * y
    cy = n[2] * x - n[0] * z
    cz = n[0] * y - n[1] * x
    xr = x * c + cx * s + n[0] * dot * (1 - c)
    yr = y * c + cy * s + n[1] * dot * (1 - c)
    zr = z * c + cz * s + n[2] * dot * (1 - c)
    return cosine_bell(xr, yr, zr, lat0, lon0, radius, h0)


def bell_centre(seed=12345):
    """Seed-chosen bell centre: latitude U(-60, 60) degrees, longitude
    U[0, 360) degrees, returned in radians."""
    rng = np.random.default_rng(seed)
    return np.deg2rad(rng.uniform(-60.0, 60.0)), np.deg2rad(rng.uniform(0.0, 360.0))
