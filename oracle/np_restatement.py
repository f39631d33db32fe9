"""Second, independent restatement of forward_comp.m in vectorised numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Written from
forward_comp.m:15-90 with numpy's own linear algebra and libm (np.arccos,
np.sin, np.arctan), so it shares no code with the C oracle or the engine.  Its
last bits can differ from theirs; the tests compare decisions and accept a
disagreement only where the unrounded pixel coordinate or a window test sits
within a tiny margin of its boundary.
"""
from __future__ import annotations

import numpy as np


def forward_comp_np(geo, proj: np.ndarray, U, pos, rot_angles, hkl):
    """Returns (Nr_simu, Nr_intersect, dis_median, events) where events is a
    dict of per-selected-event arrays plus the unrounded coordinates of every
    event (for margin checks)."""
    S = np.asarray(geo.S, float).reshape(3, 3)
    B = np.asarray(geo.B, float).reshape(3, 3)
    R = np.asarray(geo.RotDet, float).reshape(3, 3)
    U = np.asarray(U, float).reshape(3, 3)
    Lss, Lsd = geo.Lsam2sou, geo.Lsam2det
    L = Lss + Lsd
    RAO = geo.RotAxisOffset
    hkl_t = np.asarray(hkl, float).reshape(-1, 3).T  # 3 x H (forward_comp.m:15)
    Gw = S @ U @ B @ hkl_t  # :16
    pos = np.array(pos, float).reshape(1, 3).copy()
    pos[0, 1] -= RAO  # :17
    thr = geo.minEucDis / np.mean([geo.pixelysize, geo.pixelzsize])
    out_sel, raw = [], []
    dis_all = []
    Nr_simu = Nr_int = 0
    for i, ang in enumerate(np.asarray(rot_angles, float).reshape(-1)):
        om = ang * np.pi / 180
        Om = np.array([[np.cos(om), -np.sin(om), 0], [np.sin(om), np.cos(om), 0], [0, 0, 1]])
        SW = Om @ S @ pos.T  # :22
        SW = SW[:, 0]
        center = np.array([L, (SW[1] - geo.P0y + RAO) * L / (Lss + SW[0]), (SW[2] - geo.P0z) * L / (Lss + SW[0])])
        alpha = np.arctan(np.sqrt((SW[1] - geo.P0y + RAO) ** 2 + (SW[2] - geo.P0z) ** 2) / (Lss + SW[0]))
        grainpos = np.array([Lss + SW[0], SW[1] - geo.P0y + RAO, SW[2] - geo.P0z])
        Gt = Om @ Gw  # :28
        v1 = np.vstack([np.zeros(Gt.shape[1]), Gt[1], Gt[2]])
        Glen = np.sqrt(Gt[0] ** 2 + Gt[1] ** 2 + Gt[2] ** 2)
        beta = np.arccos(np.sum((grainpos / np.linalg.norm(grainpos))[:, None] * (Gt / Glen), axis=0))  # :31
        theta = beta - np.pi / 2
        sintth = np.sin(2 * theta)
        d = 1 / Glen * 2 * np.pi
        lam = 2 * d * np.sin(theta)
        with np.errstate(invalid="ignore", divide="ignore"):
            phix = np.arccos(np.sum((v1 / np.sqrt(np.sum(v1 ** 2, axis=0))) * (center / np.linalg.norm(center))[:, None],
                                    axis=0))
            phiy = phix - 2 * theta
            L2 = (Lsd - SW[0]) / np.cos(alpha)
            diffvec = L2 * sintth / np.sin(phiy)
            konst = np.sqrt(Gt[1] ** 2 + Gt[2] ** 2)
            dety22 = center[1] + diffvec * Gt[1] / konst
            detz22 = center[2] + diffvec * Gt[2] / konst
            K = np.vstack([np.full(Gt.shape[1], Lsd), dety22, detz22]) - SW[:, None]
            K = K / np.sqrt(np.sum(K ** 2, axis=0))
            t = (R[0, 0] * (Lsd - SW[0]) + R[1, 0] * (geo.dety00 - RAO - SW[1]) + R[2, 0] * (geo.detz00 - SW[2])) / (
                R[0, 0] * K[0] + R[1, 0] * K[1] + R[2, 0] * K[2])
            off = np.array([SW[0] - Lsd, SW[1] - geo.dety00 + RAO, SW[2] - geo.detz00])
            w = t * K + off[:, None]
            Y22 = R[:, 1] @ w
            Z22 = R[:, 2] @ w
            ry = -Y22 / geo.pixelysize + geo.dety0
            rz = -Z22 / geo.pixelzsize + geo.detz0
        dety = np.where(np.isfinite(ry), np.sign(ry) * np.floor(np.abs(ry) + 0.5), np.nan)
        detz = np.where(np.isfinite(rz), np.sign(rz) * np.floor(np.abs(rz) + 0.5), np.nan)
        beta_max = (90 + geo.thetamax * 4) / 180 * np.pi
        with np.errstate(invalid="ignore"):
            sel = ((beta > np.pi / 2) & (beta < beta_max) & (lam > geo.lambda_min) & (lam < geo.lambda_max)
                   & (dety >= 1) & (dety <= geo.detysize) & (detz >= 1) & (detz <= geo.detzsize)
                   & ~((dety >= geo.BeamStopY[0]) & (dety <= geo.BeamStopY[1]) & (detz >= geo.BeamStopZ[0])
                       & (detz <= geo.BeamStopZ[1])))
        raw.append(np.stack([np.full_like(beta, i), np.arange(beta.size), beta, lam, ry, rz], axis=1))
        for j in np.nonzero(sel)[0]:
            dis = proj[int(detz[j]) - 1, int(dety[j]) - 1, i]
            Nr_simu += 1
            if float(dis) <= thr:
                Nr_int += 1
            dis_all.append(np.float32(dis))
            out_sel.append((i, j, dety[j], detz[j], Y22[j], Z22[j], float(dis)))
    if dis_all:
        v = np.sort(np.array(dis_all, dtype=np.float32))
        n = v.size
        if n % 2:
            med = v[n // 2]
        else:
            a, b = v[n // 2 - 1], v[n // 2]
            med = np.float32(a + np.float32(b - a) / np.float32(2))
    else:
        med = np.float32(2000)
    return Nr_simu, Nr_int, float(med), {"sel": np.array(out_sel).reshape(-1, 7), "raw": np.concatenate(raw)}
