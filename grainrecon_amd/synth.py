"""Synthetic lab-DCT data sets for measurement and tests.

A cylindrical sample is cut into seeded Voronoi grains with Haar-uniform
orientations (Shoemake); spot masks are forward-simulated on the GPU with the
scoring geometry (every voxel of every grain projects its spots) and turned
into a bwdist-style distance stack on the GPU.  Candidate orientation sets are
seeded random samples reduced to the cubic (Oh) fundamental zone.

Voxel index -> position follows ind2pos.m:6-10 with RecVolumePixel = 1,
Center = 0 and S = I (the simap convention of the configs).
"""
from __future__ import annotations

import dataclasses
import itertools

import numpy as np

from .geometry import Geometry, cubic_hkl, labdct_geometry


def haar_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotations (K. Shoemake, Graphics Gems III, 1992), q0 >= 0."""
    u1, u2, u3 = rng.random((3, n))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    q[q[:, 0] < 0] *= -1
    return q


def cubic_symmetry_quats() -> np.ndarray:
    """The 24 proper rotations of the cube as unit quaternions (w, x, y, z):
    identity, 90/180/270 degrees about the cube axes, 180 degrees about the six
    face diagonals and 120/240 degrees about the four body diagonals."""
    h = np.sqrt(0.5)
    out = [(1.0, 0.0, 0.0, 0.0)]
    for ax in range(3):
        for w, v in ((h, h), (0.0, 1.0), (-h, h)):
            q = [w, 0.0, 0.0, 0.0]
            q[1 + ax] = v
            out.append(tuple(q))
    for a, b in ((1, 2), (1, 3), (2, 3)):
        for sgn in (1.0, -1.0):
            q = [0.0, 0.0, 0.0, 0.0]
            q[a], q[b] = h, sgn * h
            out.append(tuple(q))
    for sx, sy, sz in itertools.product((0.5, -0.5), repeat=3):
        out.append((0.5, sx, sy, sz))
    return np.array(out)


def _qmul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    w1, x1, y1, z1 = np.moveaxis(a, -1, 0)
    w2, x2, y2, z2 = np.moveaxis(b, -1, 0)
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], axis=-1)


def fz_reduce_cubic(q: np.ndarray) -> np.ndarray:
    """Symmetric equivalent closest to the identity (largest |q0|), q0 >= 0:
    the cubic-Oh fundamental zone (|r_i| <= sqrt(2)-1, sum|r_i| <= 1)."""
    sym = cubic_symmetry_quats()
    out = np.empty_like(q)
    for b in range(0, q.shape[0], 65536):
        qq = q[b:b + 65536]
        cand = _qmul(qq[:, None, :], sym[None, :, :])
        k = np.argmax(np.abs(cand[:, :, 0]), axis=1)
        r = cand[np.arange(qq.shape[0]), k]
        r[r[:, 0] < 0] *= -1
        out[b:b + 65536] = r
    return out


def quat_to_U(q: np.ndarray) -> np.ndarray:
    """quaternion2U.m:3-28 for rows of q; returns (n, 3, 3)."""
    q0, q1, q2, q3 = q.T
    return np.stack([
        np.stack([2 * (q0 * q0 + q1 * q1) - 1, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)], -1),
        np.stack([2 * (q1 * q2 + q0 * q3), 2 * (q0 * q0 + q2 * q2) - 1, 2 * (q2 * q3 - q0 * q1)], -1),
        np.stack([2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 2 * (q0 * q0 + q3 * q3) - 1], -1),
    ], 1)


@dataclasses.dataclass
class Sample:
    dims: tuple            # voxel grid (nx, ny, nz)
    vox_mm: float
    index: np.ndarray      # (M, 3) 1-based voxel indices inside the cylinder
    pos: np.ndarray        # (M, 3) positions in mm
    grain: np.ndarray      # (M,) 0-based grain id
    q: np.ndarray          # (G, 4) grain orientations
    U: np.ndarray          # (G, 3, 3)
    seeds: np.ndarray      # (G, 3) Voronoi seeds, mm


def voronoi_cylinder(diam_um: float, height_um: float, vox_um: float, n_grains: int, seed: int) -> Sample:
    """Cylinder (axis = rotation axis z) of seeded Voronoi grains."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    vox = vox_um * 1e-3
    nxy = int(round(diam_um / vox_um))
    nzh = int(round(height_um / vox_um))
    R = diam_um * 1e-3 / 2
    ii = np.arange(1, nxy + 1)
    kk = np.arange(1, nzh + 1)
    # ind2pos.m:6-10 with RecVolumePixel(:,1) = 1, Center = 0, Dimension = dims
    cxy = ((ii + 1 - 1) - nxy / 2) * vox
    cz = ((kk + 1 - 1) - nzh / 2) * vox
    X, Y = np.meshgrid(cxy, cxy, indexing="ij")
    inside = (X - vox / 2) ** 2 + (Y - vox / 2) ** 2 <= R * R
    I, J = np.nonzero(inside)
    nin = I.size
    index = np.empty((nin * nzh, 3), np.int32)
    index[:, 0] = np.repeat(ii[I], nzh)
    index[:, 1] = np.repeat(ii[J], nzh)
    index[:, 2] = np.tile(kk, nin)
    pos = np.empty((index.shape[0], 3))
    pos[:, 0] = cxy[index[:, 0] - 1]
    pos[:, 1] = cxy[index[:, 1] - 1]
    pos[:, 2] = cz[index[:, 2] - 1]
    rs = R * np.sqrt(rng.random(n_grains))
    th = 2 * np.pi * rng.random(n_grains)
    seeds = np.stack([rs * np.cos(th), rs * np.sin(th), (rng.random(n_grains) - 0.5) * height_um * 1e-3], 1)
    _, grain = cKDTree(seeds).query(pos, workers=-1)
    q = haar_quaternions(rng, n_grains)
    return Sample(dims=(nxy, nxy, nzh), vox_mm=vox, index=index, pos=pos, grain=grain.astype(np.int32), q=q,
                  U=quat_to_U(q), seeds=seeds)


@dataclasses.dataclass
class Config:
    name: str
    geo: Geometry
    hkl: np.ndarray
    rot: np.ndarray
    sample: Sample


def fe_config(scale: str = "cfg2", seed: int = 1002) -> Config:
    """The BCC Fe configurations: 'cfg1' (100 um cylinder at 5 um, 6 grains,
    1024^2 detector at 6.72 um) and 'cfg2' (400 um at 2.5 um, 300 grains,
    2048^2 at 3.36 um); Lsam2sou = Lsam2det = 14 mm, 181 omega steps of 2
    degrees, 10-70 keV, families {110} {200} {211} {220} {310} {222}."""
    if scale == "cfg1":
        geo = labdct_geometry(a=2.8665, lattice="I", det_pixels=1024, pixel_mm=6.72e-3)
        sample = voronoi_cylinder(100.0, 100.0, 5.0, 6, seed)
    elif scale == "cfg2":
        geo = labdct_geometry(a=2.8665, lattice="I", det_pixels=2048, pixel_mm=3.36e-3)
        sample = voronoi_cylinder(400.0, 400.0, 2.5, 300, seed)
    else:
        raise ValueError(scale)
    return Config(scale, geo, cubic_hkl("I", 6), np.arange(0.0, 361.0, 2.0), sample)


def al_config(seed: int = 1005) -> Config:
    """The FCC Al configuration: a = 4.0495 A, families {111} {200} {220}
    {311} {222} {400} (64 reflections), 361 omega steps of 1 degree, 2048^2
    detector at 3.36 um, Lsam2sou = Lsam2det = 14 mm, 10-70 keV; a 400 um
    cylinder at 2.5 um voxels, 310 um high (about 2.5 M voxels), 1000
    Voronoi grains."""
    geo = labdct_geometry(a=4.0495, lattice="F", det_pixels=2048, pixel_mm=3.36e-3)
    sample = voronoi_cylinder(400.0, 310.0, 2.5, 1000, seed)
    return Config("cfg5", geo, cubic_hkl("F", 6), np.arange(0.0, 361.0, 1.0), sample)


def subvolumes(sample: Sample, fractions) -> np.ndarray:
    """Splits the voxels into z slabs holding the given fractions of the
    voxels (uneven subvolumes); returns the slab index of every voxel."""
    order = np.argsort(sample.index[:, 2], kind="stable")
    bounds = np.round(np.cumsum(np.asarray(fractions, float) / np.sum(fractions)) * order.size).astype(np.int64)
    slab = np.empty(order.size, np.int32)
    start = 0
    for k, b in enumerate(bounds):
        slab[order[start:b]] = k
        start = b
    return slab


def grain_id_volume(sample: Sample) -> np.ndarray:
    """DS.GrainId-style volume: 1-based grain id inside the sample, 0 outside,
    indexed (x, y, z) by the 1-based voxel indices of the sample."""
    vol = np.zeros(sample.dims, np.int32)
    i = sample.index
    vol[i[:, 0] - 1, i[:, 1] - 1, i[:, 2] - 1] = sample.grain + 1
    return vol


def load_on_engine(eng, cfg: Config) -> None:
    """Geometry + forward-simulated spots + GPU distance transform on `eng`."""
    eng.set_geometry(cfg.geo, cfg.hkl, cfg.rot)
    eng.render_spots(cfg.sample.pos, cfg.sample.grain, cfg.sample.U.reshape(-1, 9))
    eng.distance_from_mask()
