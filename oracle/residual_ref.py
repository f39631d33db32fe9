"""Reference restatement of the spot-centroid residual in torch float64 -- TEST
INFRASTRUCTURE ONLY (gradients by torch autograd pin the engine's hand-written
adjoint).

The residual of one matched spot is the difference between the simulated and
the measured centroid, in pixels, before rounding:
    r = (-dety22/pixelysize + dety0 - y_meas, -detz22/pixelzsize + detz0 - z_meas)
with dety22/detz22 from the forward_comp geometry (forward_comp.m:15-55, the
convention of dis_calc_spotspair.m:35-71 and fit_grain_COM.m:49-83).
Parameters: the 7 geometry values x = [Lsam2sou Lsam2det dety00 detz00 tilt_x
tilt_y tilt_z] (geo_fitting.m:8-17, tilts in degrees through get_det_R.m) and
per grain its Rodrigues vector r (U = quaternion2U(rod2quat(r)), the path of
GrainMapping_engine_gpu_comp.m:151-152,200) and centre of mass (mm).
cos(alpha) is used in its algebraic form gp1/|gp| (alpha = atan(sqrt(gp2^2 +
gp3^2)/gp1), forward_comp.m:25), which is smooth at gp2 = gp3 = 0.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _rot_det(t):
    """get_det_R.m: RotZ*RotY*RotX with angles in degrees."""
    a = t * (math.pi / 180.0)
    cx, sx = torch.cos(a[0]), torch.sin(a[0])
    cy, sy = torch.cos(a[1]), torch.sin(a[1])
    cz, sz = torch.cos(a[2]), torch.sin(a[2])
    one, zero = torch.ones((), dtype=t.dtype), torch.zeros((), dtype=t.dtype)
    RX = torch.stack([torch.stack([one, zero, zero]), torch.stack([zero, cx, -sx]), torch.stack([zero, sx, cx])])
    RY = torch.stack([torch.stack([cy, zero, sy]), torch.stack([zero, one, zero]), torch.stack([-sy, zero, cy])])
    RZ = torch.stack([torch.stack([cz, -sz, zero]), torch.stack([sz, cz, zero]), torch.stack([zero, zero, one])])
    return RZ @ RY @ RX


def _U_from_rod(r):
    """quaternion2U(rod2quat(r)) for rows of r (n, 3)."""
    s = 1.0 / torch.sqrt(1.0 + (r * r).sum(-1))
    q0, q1, q2, q3 = s, r[:, 0] * s, r[:, 1] * s, r[:, 2] * s
    return torch.stack([
        torch.stack([2 * (q0 * q0 + q1 * q1) - 1, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)], -1),
        torch.stack([2 * (q1 * q2 + q0 * q3), 2 * (q0 * q0 + q2 * q2) - 1, 2 * (q2 * q3 - q0 * q1)], -1),
        torch.stack([2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 2 * (q0 * q0 + q3 * q3) - 1], -1),
    ], 1)


def residual(geo, spots, g7, grains6):
    """geo: object with S, B, P0y, P0z, RotAxisOffset, pixelysize, pixelzsize,
    dety0, detz0.  spots: dict of numpy arrays hkl (n,3), omega (n,) degrees,
    meas (n,2) pixels, grain (n,) int.  g7: tensor (7,); grains6: tensor (G,6)
    [r1 r2 r3 x y z].  Returns (n, 2) residuals."""
    dt = torch.float64
    S = torch.tensor(np.asarray(geo.S, float), dtype=dt)
    B = torch.tensor(np.asarray(geo.B, float), dtype=dt)
    hkl = torch.tensor(spots["hkl"], dtype=dt)
    om = torch.tensor(spots["omega"], dtype=dt) * (math.pi / 180.0)
    meas = torch.tensor(spots["meas"], dtype=dt)
    gi = torch.tensor(spots["grain"], dtype=torch.long)
    Lss, Lsd, dy00, dz00 = g7[0], g7[1], g7[2], g7[3]
    R = _rot_det(g7[4:7])
    RAO, P0y, P0z = geo.RotAxisOffset, geo.P0y, geo.P0z
    U = _U_from_rod(grains6[:, :3])[gi]
    P = grains6[:, 3:][gi]
    c, s = torch.cos(om), torch.sin(om)
    zero, one = torch.zeros_like(c), torch.ones_like(c)
    Om = torch.stack([torch.stack([c, -s, zero], -1), torch.stack([s, c, zero], -1),
                      torch.stack([zero, zero, one], -1)], 1)
    Gw = (S @ U @ B @ hkl.unsqueeze(-1)).squeeze(-1)
    pos = P.clone()
    pos = torch.stack([pos[:, 0], pos[:, 1] - RAO, pos[:, 2]], -1)
    SW = (Om @ S @ pos.unsqueeze(-1)).squeeze(-1)
    L = Lss + Lsd
    gp = torch.stack([Lss + SW[:, 0], SW[:, 1] - P0y + RAO, SW[:, 2] - P0z], -1)
    ctr = torch.stack([L * torch.ones_like(c), gp[:, 1] * L / gp[:, 0], gp[:, 2] * L / gp[:, 0]], -1)
    ng = torch.sqrt((gp * gp).sum(-1))
    u = gp / ng.unsqueeze(-1)
    cu = ctr / torch.sqrt((ctr * ctr).sum(-1)).unsqueeze(-1)
    L2 = (Lsd - SW[:, 0]) * ng / gp[:, 0]
    Gt = (Om @ Gw.unsqueeze(-1)).squeeze(-1)
    Glen = torch.sqrt((Gt * Gt).sum(-1))
    beta = torch.acos((u * Gt).sum(-1) / Glen)
    theta = beta - math.pi / 2
    nv = torch.sqrt(Gt[:, 1] ** 2 + Gt[:, 2] ** 2)
    phix = torch.acos((Gt[:, 1] * cu[:, 1] + Gt[:, 2] * cu[:, 2]) / nv)
    phiy = phix - 2 * theta
    diffvec = L2 * torch.sin(2 * theta) / torch.sin(phiy)
    y22 = ctr[:, 1] + diffvec * Gt[:, 1] / nv
    z22 = ctr[:, 2] + diffvec * Gt[:, 2] / nv
    k = torch.stack([Lsd - SW[:, 0], y22 - SW[:, 1], z22 - SW[:, 2]], -1)
    K = k / torch.sqrt((k * k).sum(-1)).unsqueeze(-1)
    tnum = R[0, 0] * (Lsd - SW[:, 0]) + R[1, 0] * (dy00 - RAO - SW[:, 1]) + R[2, 0] * (dz00 - SW[:, 2])
    tden = R[0, 0] * K[:, 0] + R[1, 0] * K[:, 1] + R[2, 0] * K[:, 2]
    t = tnum / tden
    w = t.unsqueeze(-1) * K + torch.stack([SW[:, 0] - Lsd, SW[:, 1] - dy00 + RAO, SW[:, 2] - dz00], -1)
    Y22 = R[0, 1] * w[:, 0] + R[1, 1] * w[:, 1] + R[2, 1] * w[:, 2]
    Z22 = R[0, 2] * w[:, 0] + R[1, 2] * w[:, 1] + R[2, 2] * w[:, 2]
    ry = -Y22 / geo.pixelysize + geo.dety0
    rz = -Z22 / geo.pixelzsize + geo.detz0
    return torch.stack([ry - meas[:, 0], rz - meas[:, 1]], -1)


def loss_and_grad(geo, spots, g7, grains6):
    """0.5*sum(r^2) and its gradient with respect to (g7, grains6), by autograd."""
    a = torch.tensor(np.asarray(g7, float), dtype=torch.float64, requires_grad=True)
    b = torch.tensor(np.asarray(grains6, float), dtype=torch.float64, requires_grad=True)
    r = residual(geo, spots, a, b)
    f = 0.5 * (r * r).sum()
    f.backward()
    return float(f), a.grad.numpy().copy(), b.grad.numpy().copy(), r.detach().numpy()
