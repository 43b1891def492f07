"""Synthetic particle configurations of the benchmark workloads (seeded, numpy).

All lengths are in sphere radii (a = 1); the no-slip wall is the plane z = 0.
Segment positions are returned in the reference's AoS layout, filament after
filament, base segment first (the ordering of x_segs_host,
src/cilia/cilia_sim_main.cu:274-279).

* ``tilted_lattice``  -- 64 filaments x 20 segments, 8x8 lattice (spacing 10),
  bases at z = 1.1, straight chains of spacing 2.2 tilted by a uniform polar
  angle <= 0.3 rad and a uniform azimuth (1,280 segments).
* ``wave_lattice``    -- 640 filaments x 20 segments on a 32x20 lattice (spacing
  8), planar bends of tip amplitude 0.2 L whose phase is 2 pi x / lambda with
  lambda = 8 lattice spacings (12,800 segments).
* ``sphere_over_wall`` -- a rigid sphere of radius 40 carrying 6,000 surface blobs
  (Fibonacci lattice) and 1,300 radial filaments x 20 segments seeded on a second
  Fibonacci lattice, bent by a metachronal phase field; the lowest point of the
  sphere is 50 above the wall (about 32,000 particles).
* ``ensemble``        -- independent wall systems with segment counts uniform in
  [100, 4000] (multiples of 20), each a scaled ``tilted_lattice``.
"""
from __future__ import annotations

import math

import numpy as np

DL = 2.2     # segment spacing used by every workload
NSEG = 20


def _chain(base, tangents, dl=DL):
    """Segment centres from a base point and per-segment unit tangents (robot arm,
    src/cilia/filament.cpp:629-651)."""
    x = np.empty((len(tangents), 3))
    x[0] = base
    for k in range(1, len(tangents)):
        x[k] = x[k - 1] + 0.5 * dl * (tangents[k - 1] + tangents[k])
    return x


def loads(n, dof, seed):
    """i.i.d. N(0,1) forces (and torques) in AoS layout."""
    return np.random.default_rng(seed).standard_normal(dof * n)


def tilted_lattice(nx=8, ny=8, spacing=10.0, nseg=NSEG, z0=1.1, max_polar=0.3, seed=0,
                   return_tangents=False, min_sep=None, count=None):
    """Straight filaments on an nx x ny lattice, each tilted by a random polar angle
    in [0, max_polar) and azimuth in [0, 2 pi).  With min_sep, a filament's tilt is
    redrawn until none of its segments is closer than min_sep to a segment of an
    earlier filament (tilted neighbours can otherwise cross).  count keeps only the
    first `count` lattice sites."""
    rng = np.random.default_rng(seed)
    fils, tans = [], []
    sites = [(ix, iy) for ix in range(nx) for iy in range(ny)]
    if count is not None:
        sites = sites[:count]
    for ix, iy in sites:
        base = np.array([ix * spacing, iy * spacing, z0])
        placed = np.concatenate(fils) if (min_sep and fils) else None
        for attempt in range(1000):
            th = rng.uniform(0.0, max_polar)
            ph = rng.uniform(0.0, 2.0 * math.pi)
            t = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
            tt = np.repeat(t[None], nseg, axis=0)
            chain = _chain(base, tt)
            if placed is None:
                break
            near = placed[np.abs(placed[:, 0] - base[0]) + np.abs(placed[:, 1] - base[1])
                          < 2.0 * nseg * DL]
            if near.size == 0:
                break
            d2 = ((chain[:, None, :] - near[None, :, :]) ** 2).sum(-1)
            if d2.min() >= min_sep * min_sep:
                break
        else:
            raise RuntimeError("could not place a non-overlapping filament")
        fils.append(chain)
        tans.append(tt)
    x = np.concatenate(fils).reshape(-1)
    return (x, np.concatenate(tans)) if return_tangents else x


def _bent_tangents(nseg, amp_frac, phase, axis_u, axis_n, dl=DL):
    """Tangents of a planar bend in the (axis_u, axis_n) plane whose tip is
    displaced by about amp_frac * L * cos(phase) along axis_n (L = nseg * dl)."""
    L = nseg * dl
    s = (np.arange(nseg) + 0.5) * dl
    theta = 2.0 * amp_frac * (s / L) * math.cos(phase)
    return np.cos(theta)[:, None] * axis_u[None] + np.sin(theta)[:, None] * axis_n[None]


def wave_lattice(nx=32, ny=20, spacing=8.0, nseg=NSEG, z0=1.1, amp_frac=0.2, wavelength=None,
                 return_tangents=False):
    wavelength = 8.0 * spacing if wavelength is None else wavelength
    ez, ex = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0])
    fils, tans = [], []
    for ix in range(nx):
        for iy in range(ny):
            base = np.array([ix * spacing, iy * spacing, z0])
            phase = 2.0 * math.pi * base[0] / wavelength
            tt = _bent_tangents(nseg, amp_frac, phase, ez, ex)
            fils.append(_chain(base, tt))
            tans.append(tt)
    x = np.concatenate(fils).reshape(-1)
    return (x, np.concatenate(tans)) if return_tangents else x


def fibonacci_sphere(n, radius=1.0, offset=0.0):
    """n nearly uniform points on a sphere (golden-angle spiral)."""
    k = np.arange(n) + 0.5
    z = 1.0 - 2.0 * k / n
    r = np.sqrt(1.0 - z * z)
    phi = math.pi * (3.0 - math.sqrt(5.0)) * k + offset
    return radius * np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def sphere_over_wall(n_blob=6000, n_fil=1300, nseg=NSEG, radius=40.0, gap=50.0,
                     amp_frac=0.08, wavelength=None, return_tangents=False):
    """Returns (x_seg AoS, x_blob AoS, centre, filament base directions[, tangents])."""
    centre = np.array([0.0, 0.0, gap + radius])
    blobs = fibonacci_sphere(n_blob, radius) + centre
    dirs = fibonacci_sphere(n_fil, 1.0, offset=0.5)
    wavelength = 0.5 * math.pi * radius if wavelength is None else wavelength
    fils, tans = [], []
    for d in dirs:
        # beat plane spanned by the normal and the local polar direction
        pol = np.array([0.0, 0.0, -1.0]) + d[2] * d   # tangential part of -z
        nrm = np.linalg.norm(pol)
        pol = pol / nrm if nrm > 1e-12 else np.array([1.0, 0.0, 0.0])
        polar_angle = math.acos(max(-1.0, min(1.0, d[2])))
        phase = 2.0 * math.pi * radius * polar_angle / wavelength   # metachronal wave
        base = centre + (radius + DL) * d
        tt = _bent_tangents(nseg, amp_frac, phase, d, pol)
        fils.append(_chain(base, tt))
        tans.append(tt)
    out = (np.concatenate(fils).reshape(-1), blobs.reshape(-1), centre, dirs)
    return out + (np.concatenate(tans),) if return_tangents else out


def ensemble(n_sys=64, lo=100, hi=4000, seed=3):
    """List of (n_seg, x_seg) for independent wall systems of uneven size: lattices of
    tilted filaments as tilted_lattice, with no two segments of different filaments
    closer than one segment spacing."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_sys):
        nfil = int(rng.integers(lo // NSEG, hi // NSEG + 1))
        side = int(math.ceil(math.sqrt(nfil)))
        x = tilted_lattice(nx=side, ny=side, seed=1000 + s, min_sep=2.2, count=nfil)
        out.append((nfil * NSEG, x))
    return out


def min_separation(x):
    """Smallest centre distance (brute force; small systems only)."""
    p = np.asarray(x).reshape(-1, 3)
    d = np.sqrt(((p[:, None, :] - p[None, :, :]) ** 2).sum(-1))
    np.fill_diagonal(d, np.inf)
    return d.min()


def _quat_mul(p, q):
    return np.array([p[0] * q[0] - p[1:] @ q[1:],
                     *(p[0] * q[1:] + q[0] * p[1:] + np.cross(p[1:], q[1:]))])


def quat_from_x_to(t):
    """Unit quaternion (s, v) of the rotation taking the x axis to the unit vector t
    (the construction of filament::initial_setup, filament.cpp: qtemp.sqrt_in_place_ortho)."""
    t = np.asarray(t, float) / np.linalg.norm(t)
    q = np.array([t[0], 0.0, -t[2], t[1]])
    if q[0] < -0.999999:
        q[0] = -0.9999999
    s = math.sqrt(0.5 * (1.0 + q[0]))
    q = np.concatenate([[s], q[1:] / (2.0 * s)])
    return q / np.linalg.norm(q)


def segment_quaternions(tangents, nseg=NSEG):
    """Segment orientations whose first frame vector is the given tangent: the base
    from quat_from_x_to, each next one by the minimal rotation between successive
    tangents (no twist), AoS (s, v0, v1, v2) per segment."""
    T = np.asarray(tangents, float).reshape(-1, nseg, 3)
    out = np.empty((T.shape[0], nseg, 4))
    for f in range(T.shape[0]):
        q = quat_from_x_to(T[f, 0])
        out[f, 0] = q
        for k in range(1, nseg):
            a = T[f, k - 1] / np.linalg.norm(T[f, k - 1])
            b = T[f, k] / np.linalg.norm(T[f, k])
            r = np.concatenate([[1.0 + a @ b], np.cross(a, b)])
            r /= np.linalg.norm(r)
            q = _quat_mul(r, q)
            q /= np.linalg.norm(q)
            out[f, k] = q
    return out.reshape(-1)
