"""Geometry and crystallography inputs of the completeness scorer.

Host-side input preparation (run once per reconstruction), restated from
GrainRecon:
  form_B          FormB.m:15-51            reciprocal-lattice matrix B (2*pi convention)
  get_det_R       get_det_R.m:4-7          detector rotation RotZ*RotY*RotX, degrees
  cubic_hkl       setup_hkl.m:46-63, genhkl.m, sysabs.m
                                           reflections of the first n families with
                                           Friedel mates, systematic absences removed
  theta_max_deg   setup_hkl.m:17-20        half of the detector-corner two-theta
  lambda_window   setup_hkl.m:21-22        12.39818746/[Emax, Emin] Angstrom
  euler2u         euler2u.m:5-16
  quaternion2U    quaternion2U.m:3-28
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Sequence

import numpy as np

KEV_ANGSTROM = 12.39818746  # hc in keV*Angstrom (setup_hkl.m:21)


def form_B(cell: Sequence[float]) -> np.ndarray:
    """FormB.m:15-51: B with the physicist 2*pi convention."""
    a, b, c, al, be, ga = [float(x) for x in cell]
    calp, cbet, cgam = (math.cos(x * math.pi / 180.0) for x in (al, be, ga))
    salp, sbet, sgam = (math.sin(x * math.pi / 180.0) for x in (al, be, ga))
    angular = math.sqrt(1 - calp**2 - cbet**2 - cgam**2 + 2 * calp * cbet * cgam)
    V = a * b * c * angular
    astar = 2 * math.pi * b * c * salp / V
    bstar = 2 * math.pi * a * c * sbet / V
    cstar = 2 * math.pi * a * b * sgam / V
    sbetstar = V / (a * b * c * salp * sgam)
    sgamstar = V / (a * b * c * salp * sbet)
    cbetstar = (calp * cgam - cbet) / (salp * sgam)
    cgamstar = (calp * cbet - cgam) / (salp * sbet)
    B = np.zeros((3, 3))
    B[0, 0] = astar
    B[0, 1] = bstar * cgamstar
    B[0, 2] = cstar * cbetstar
    B[1, 1] = bstar * sgamstar
    B[1, 2] = -cstar * sbetstar * calp
    B[2, 2] = cstar * sbetstar * salp
    return B


def _cosd(x: float) -> float:
    # MATLAB cosd reduces in degrees so multiples of 90 give exact zeros
    r = math.fmod(x, 360.0)
    if r % 90.0 == 0.0:
        return [1.0, 0.0, -1.0, 0.0][int(round(r / 90.0)) % 4]
    return math.cos(math.radians(x))


def _sind(x: float) -> float:
    r = math.fmod(x, 360.0)
    if r % 90.0 == 0.0:
        return [0.0, 1.0, 0.0, -1.0][int(round(r / 90.0)) % 4]
    return math.sin(math.radians(x))


def get_det_R(tilt_x: float, tilt_y: float, tilt_z: float) -> np.ndarray:
    """get_det_R.m:4-7: RotDet = RotZ*RotY*RotX (angles in degrees)."""
    cx, sx = _cosd(tilt_x), _sind(tilt_x)
    cy, sy = _cosd(tilt_y), _sind(tilt_y)
    cz, sz = _cosd(tilt_z), _sind(tilt_z)
    RotX = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]], dtype=float)
    RotY = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]], dtype=float)
    RotZ = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]], dtype=float)
    return RotZ @ RotY @ RotX


def _allowed(h: int, k: int, l: int, lattice: str) -> bool:
    if lattice == "I":  # body centred: h+k+l even (sysabs.m condition 5, X=2)
        return (h + k + l) % 2 == 0
    if lattice == "F":  # face centred: h+k, h+l, k+l even (sysabs.m condition 4)
        return (h + k) % 2 == 0 and (h + l) % 2 == 0 and (k + l) % 2 == 0
    return True


def cubic_hkl(lattice: str, n_families: int, hmax: int = 8) -> np.ndarray:
    """All signed reflections (Friedel mates included) of the first
    `n_families` allowed {hkl} families of a cubic lattice, ordered by
    h^2+k^2+l^2 (setup_hkl.m:46-63 keeps families up to the n-th value of
    h^2+k^2+l^2).  Returns an (H, 3) float array."""
    refl = []
    for h in range(-hmax, hmax + 1):
        for k in range(-hmax, hmax + 1):
            for l in range(-hmax, hmax + 1):
                if (h, k, l) == (0, 0, 0) or not _allowed(h, k, l, lattice):
                    continue
                refl.append((h * h + k * k + l * l, h, k, l))
    sq = sorted({r[0] for r in refl})[:n_families]
    keep = sorted([r for r in refl if r[0] in sq], key=lambda r: (r[0], -r[1], -r[2], -r[3]))
    return np.array([[r[1], r[2], r[3]] for r in keep], dtype=float)


def theta_max_deg(Lsam2det: float, detysize: float, detzsize: float, pixelysize: float, pixelzsize: float) -> float:
    """setup_hkl.m:17-20: half the two-theta of the detector corner."""
    a = np.array([Lsam2det, 0.0, 0.0])
    b = np.array([Lsam2det, 0.5 * detysize * pixelysize, 0.5 * detzsize * pixelzsize])
    tth = math.acos(float(np.dot(a / np.linalg.norm(a), b / np.linalg.norm(b))))
    return tth * 180.0 / math.pi / 2.0


def lambda_window(e_min_kev: float, e_max_kev: float) -> tuple[float, float]:
    """setup_hkl.m:21-22: (lambda_min, lambda_max) in Angstrom."""
    return KEV_ANGSTROM / e_max_kev, KEV_ANGSTROM / e_min_kev


def euler2u(phi1: float, PHI: float, phi2: float) -> np.ndarray:
    """euler2u.m:5-16 (radians)."""
    c1, s1, c2, s2, cP, sP = (math.cos(phi1), math.sin(phi1), math.cos(phi2), math.sin(phi2), math.cos(PHI),
                              math.sin(PHI))
    return np.array([
        [c1 * c2 - s1 * s2 * cP, -c1 * s2 - s1 * c2 * cP, s1 * sP],
        [s1 * c2 + c1 * s2 * cP, -s1 * s2 + c1 * c2 * cP, -c1 * sP],
        [s2 * sP, c2 * sP, cP],
    ])


def quaternion2U(q: Sequence[float]) -> np.ndarray:
    """quaternion2U.m:3-28."""
    q0, q1, q2, q3 = [float(x) for x in q]
    return np.array([
        [2 * (q0 * q0 + q1 * q1) - 1, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)],
        [2 * (q1 * q2 + q0 * q3), 2 * (q0 * q0 + q2 * q2) - 1, 2 * (q2 * q3 - q0 * q1)],
        [2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 2 * (q0 * q0 + q3 * q3) - 1],
    ])


def rod2quat(r: Sequence[float]) -> np.ndarray:
    """Rodrigues vector -> unit quaternion (1, r)/sqrt(1+|r|^2) (rod2quat.m)."""
    r = np.asarray(r, dtype=float)
    return np.concatenate([[1.0], r]) / math.sqrt(1.0 + float(r @ r))


class CGeometry(ctypes.Structure):
    """ctypes image of grs_geometry (include/grscore.h)."""
    _fields_ = [
        ("S", ctypes.c_double * 9),
        ("B", ctypes.c_double * 9),
        ("RotDet", ctypes.c_double * 9),
        ("Lsam2sou", ctypes.c_double),
        ("Lsam2det", ctypes.c_double),
        ("P0y", ctypes.c_double),
        ("P0z", ctypes.c_double),
        ("RotAxisOffset", ctypes.c_double),
        ("dety00", ctypes.c_double),
        ("detz00", ctypes.c_double),
        ("pixelysize", ctypes.c_double),
        ("pixelzsize", ctypes.c_double),
        ("dety0", ctypes.c_double),
        ("detz0", ctypes.c_double),
        ("detysize", ctypes.c_double),
        ("detzsize", ctypes.c_double),
        ("BeamStopY", ctypes.c_double * 2),
        ("BeamStopZ", ctypes.c_double * 2),
        ("thetamax", ctypes.c_double),
        ("lambda_min", ctypes.c_double),
        ("lambda_max", ctypes.c_double),
        ("minEucDis", ctypes.c_double),
    ]


@dataclasses.dataclass
class Geometry:
    """The argument list of forward_comp.m:4-6 minus the per-call inputs."""
    S: np.ndarray
    B: np.ndarray
    RotDet: np.ndarray
    Lsam2sou: float
    Lsam2det: float
    P0y: float
    P0z: float
    RotAxisOffset: float
    dety00: float
    detz00: float
    pixelysize: float
    pixelzsize: float
    dety0: float
    detz0: float
    detysize: float
    detzsize: float
    BeamStopY: tuple
    BeamStopZ: tuple
    thetamax: float
    lambda_min: float
    lambda_max: float
    minEucDis: float

    def to_c(self) -> CGeometry:
        g = CGeometry()
        for name in ("S", "B", "RotDet"):
            arr = np.ascontiguousarray(np.asarray(getattr(self, name), dtype=float).reshape(9))
            getattr(g, name)[:] = arr.tolist()
        for f in ("Lsam2sou", "Lsam2det", "P0y", "P0z", "RotAxisOffset", "dety00", "detz00", "pixelysize",
                  "pixelzsize", "dety0", "detz0", "detysize", "detzsize", "thetamax", "lambda_min", "lambda_max",
                  "minEucDis"):
            setattr(g, f, float(getattr(self, f)))
        g.BeamStopY[:] = [float(x) for x in self.BeamStopY]
        g.BeamStopZ[:] = [float(x) for x in self.BeamStopZ]
        return g

    def legacy_param(self, nr_hkl: int, nr_rot: int) -> np.ndarray:
        """The 20-value param row vector of ForwardCalc_OR_space_comp_cuda.m:25-27."""
        return np.array([nr_hkl, nr_rot, self.Lsam2sou, self.Lsam2det, self.P0y, self.P0z, self.dety00,
                         self.detz00, self.pixelysize, self.pixelzsize, self.detysize, self.detzsize,
                         self.BeamStopY[0], self.BeamStopY[1], self.BeamStopZ[0], self.BeamStopZ[1],
                         self.thetamax, self.lambda_min, self.lambda_max, self.RotAxisOffset], dtype=float)


def labdct_geometry(*, a: float, lattice: str, Lsam2sou: float = 14.0, Lsam2det: float = 14.0,
                    det_pixels: int = 1024, pixel_mm: float = 6.72e-3, energy_kev=(10.0, 70.0),
                    tilts=(0.0, 0.0, 0.0), dety00: float = 0.0, detz00: float = 0.0,
                    beamstop=((0.0, -1.0), (0.0, -1.0))) -> Geometry:
    """A lab-DCT geometry with the synthetic conventions used across tests and
    benchmarks: S = I (simap), P0 = RotAxisOffset = 0, beam centre at
    detsize/2 + 0.5, hit threshold 0.5 pixel, empty beamstop by default."""
    B = form_B([a, a, a, 90.0, 90.0, 90.0])
    lam_min, lam_max = lambda_window(*energy_kev)
    n = float(det_pixels)
    return Geometry(
        S=np.eye(3), B=B, RotDet=get_det_R(*tilts), Lsam2sou=Lsam2sou, Lsam2det=Lsam2det, P0y=0.0, P0z=0.0,
        RotAxisOffset=0.0, dety00=dety00, detz00=detz00, pixelysize=pixel_mm, pixelzsize=pixel_mm,
        dety0=n / 2 + 0.5, detz0=n / 2 + 0.5, detysize=n, detzsize=n, BeamStopY=beamstop[0],
        BeamStopZ=beamstop[1], thetamax=theta_max_deg(Lsam2det, n, n, pixel_mm, pixel_mm),
        lambda_min=lam_min, lambda_max=lam_max, minEucDis=0.5 * pixel_mm)
