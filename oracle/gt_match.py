"""G-vector matching score ("Gt" indexing method) -- TEST INFRASTRUCTURE ONLY
(checker of grs_forward_gt).

The CUDA source of cuda_forward_Gt is not in the reference.  Its CPU path,
ForwardCalc_OR_space_Gt.m:6-113, calls calc_Gt_match (the same file,
:117-205), and fit_OR_constrain_with_local_Gt.m:48,105 call a sibling,
calc_Gt_match_fun.m:1-152.  Both are restated (semantics="match" / "fun"):
  detector tilt folded into an effective distance and offsets (:129-133 /
  :12-16);
  calc_Gt_match only: pos(:,2) shifted by -RotAxisOffset (:135);
  per projection i (frame i <-> rot_angles(i)): SamposW = Omega*S*pos',
  Kin = grainpos/|grainpos|, for every spot
  (Spots{i}(:,9:10)) Kout_tr from the pixel position (:35-48),
  2theta = acosd(Kin.Kout_tr), lambda_j = 2 d_j sind(2theta/2),
  Gt_spot_j = (2 pi/lambda_j)(Kout_tr - Kin), normalised (:50-59, :88);
  NrSpotExpect = the forward_comp selection count with the centre at the
  effective distance (calcExpectSpotNr.m:6-47);
  for every reflection of the first four families (:96-149) the smallest
  angle acos(n.g)*180/pi to the spots' unit vectors of its family; a match
  when it is below 1 degree and |Gt| is within 0.1 of 2 pi/d (:100); the
  matched spot is the first spot of the projection with that angle (min's
  index, ForwardCalc_OR_space_Gt.m:225).
  Output per orientation:
    "match": [Nr_match, Gt_matched_cost, Nr_match/NrSpotExpectAll,
      mean(angles)], Gt_matched_cost = (sum_k (n_k - w_k)^2 summed over the
      matches per component k, the three sums added in order) / Nr_match
      with n the unit Gt and w the matched spot's unit vector (:195-198),
      100 without a match (:203) (ForwardCalc_OR_space_Gt.m:72);
    "fun": [Nr_match, median(angles), Nr_match/NrSpotExpectAll,
      mean(angles)] (calc_Gt_match_fun.m:70-78).
Choices the reference leaves open (parity unpinned at the last bit):
  sind(x) = sin((pi/180)*x) after reduction to |x| <= 45 by multiples of 90
  (the documented behaviour of MATLAB's sind); acos of a dot product above 1
  is taken as its real part 0; min ignores NaN; the median of an even count
  is a + (b - a)/2; sums and norms of 3-vectors are sequential; the mean and
  the cost sums are sequential sums over the matches in (projection,
  reflection) order.
"""
from __future__ import annotations

import math

import numpy as np

from . import math_eval

PI = float.fromhex("0x1.921fb54442d18p+1")


def _sin(x):
    return math_eval(0, np.asarray(x, dtype=float))


def _cos(x):
    return math_eval(1, np.asarray(x, dtype=float))


def _acos(x):
    return math_eval(2, np.asarray(x, dtype=float))


def _atan(x):
    return math_eval(3, np.asarray(x, dtype=float))


def _mround(x):
    x = np.asarray(x, dtype=float)
    return np.where(x >= 0, np.floor(x + 0.5), -np.floor(-x + 0.5))


def sind(x):
    x = np.asarray(x, dtype=float)
    n = _mround(x / 90.0)
    r = x - n * 90.0
    m = np.mod(n, 4.0)
    a = (PI / 180.0) * r
    s, c = _sin(a), _cos(a)
    return np.where(m == 0, s, np.where(m == 1, c, np.where(m == 2, -s, -c)))


def _U(q):
    q1, q2, q3 = q[1], q[2], q[3]
    ss = (q1 * q1 + q2 * q2) + q3 * q3
    q0 = math.sqrt(1.0 - ss)
    return np.array([
        [2.0 * (q0 * q0 + q1 * q1) - 1.0, 2.0 * (q1 * q2 - q0 * q3), 2.0 * (q1 * q3 + q0 * q2)],
        [2.0 * (q1 * q2 + q0 * q3), 2.0 * (q0 * q0 + q2 * q2) - 1.0, 2.0 * (q2 * q3 - q0 * q1)],
        [2.0 * (q1 * q3 - q0 * q2), 2.0 * (q2 * q3 + q0 * q1), 2.0 * (q0 * q0 + q3 * q3) - 1.0]])


def _mm(A, B):
    """3x3 product, sequential sums."""
    C = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            C[i, j] = (A[i, 0] * B[0, j] + A[i, 1] * B[1, j]) + A[i, 2] * B[2, j]
    return C


def gt_score(geo, pos, q_rows, rot, hkl, frames, fam_sq, d_possible, semantics="match", details=False):
    """frames: per projection (n_i, >=2) arrays [spot y, spot z, ...];
    semantics "match" (calc_Gt_match) or "fun" (calc_Gt_match_fun).  With
    details, also the matched angles of every orientation, in match order."""
    if semantics not in ("match", "fun"):
        raise ValueError(semantics)
    S, B, R = np.asarray(geo.S, float), np.asarray(geo.B, float), np.asarray(geo.RotDet, float)
    hkl = np.asarray(hkl, float).reshape(-1, 3)
    H = hkl.shape[0]
    Lss, Lsd, RAO = geo.Lsam2sou, geo.Lsam2det, geo.RotAxisOffset
    py, pz = geo.pixelysize, geo.pixelzsize
    fam_sq = [float(v) for v in fam_sq][:4]
    d_possible = [float(v) for v in d_possible][:4]
    glen_possible = [(2.0 * PI) / d for d in d_possible]
    # :12-16
    v = [Lsd, geo.dety00 - RAO, geo.detz00]
    d_tr = [(R[0, r] * v[0] + R[1, r] * v[1]) + R[2, r] * v[2] for r in range(3)]
    Lsd_tr = d_tr[0] / R[0, 0]
    dety00_tr = R[0, 1] * Lsd_tr - d_tr[1]
    detz00_tr = R[0, 2] * Lsd_tr - d_tr[2]
    L_tr = Lss + Lsd_tr
    beta_max = ((90.0 + geo.thetamax * 4.0) / 180.0) * PI
    fam = np.full(H, -1)
    for h in range(H):
        s2 = (hkl[h, 0] * hkl[h, 0] + hkl[h, 1] * hkl[h, 1]) + hkl[h, 2] * hkl[h, 2]
        for k, f in enumerate(fam_sq):
            if s2 == f:
                fam[h] = k
                break
    p = np.array(pos, float).reshape(3)
    if semantics == "match":
        p[1] = p[1] - RAO  # ForwardCalc_OR_space_Gt.m:135
    # per projection: spot unit vectors and the per-projection terms
    proj = []
    for i, om_deg in enumerate(np.asarray(rot, float)):
        om = (om_deg * PI) / 180.0
        c, s = float(_cos(om)[0]), float(_sin(om)[0])
        OS = np.array([[c * S[0, j] - s * S[1, j] for j in range(3)], [s * S[0, j] + c * S[1, j] for j in range(3)],
                       [S[2, j] for j in range(3)]])
        SW = [(OS[r, 0] * p[0] + OS[r, 1] * p[1]) + OS[r, 2] * p[2] for r in range(3)]
        gp = [Lss + SW[0], (SW[1] - geo.P0y) + RAO, SW[2] - geo.P0z]
        ng = math.sqrt((gp[0] * gp[0] + gp[1] * gp[1]) + gp[2] * gp[2])
        kin = [gp[0] / ng, gp[1] / ng, gp[2] / ng]
        sp = np.asarray(frames[i], float).reshape(-1, 3)
        dety, detz = sp[:, 0], sp[:, 1]
        y22 = (geo.dety0 - dety) * py
        z22 = (geo.detz0 - detz) * pz
        a = np.full(dety.shape, Lsd_tr - SW[0])
        b = (y22 - dety00_tr) - SW[1]
        cc = (z22 - detz00_tr) - SW[2]
        nk = np.sqrt((a * a + b * b) + cc * cc)
        ko = [a / nk, b / nk, cc / nk]
        tt = (_acos((kin[0] * ko[0] + kin[1] * ko[1]) + kin[2] * ko[2]) * 180.0) / PI
        half = sind(tt / 2.0)
        fac = [ko[0] - kin[0], ko[1] - kin[1], ko[2] - kin[2]]
        units = []
        for d in d_possible:
            lam = (2.0 * d) * half
            sc = (2.0 * PI) / lam
            g = [sc * fac[0], sc * fac[1], sc * fac[2]]
            n = np.sqrt((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2])
            units.append([g[0] / n, g[1] / n, g[2] / n])
        # calcExpectSpotNr terms: centre at the effective distance
        ctr1, ctr2 = (gp[1] * L_tr) / gp[0], (gp[2] * L_tr) / gp[0]
        alpha = float(_atan(math.sqrt(gp[1] * gp[1] + gp[2] * gp[2]) / gp[0])[0])
        nc = math.sqrt((L_tr * L_tr + ctr1 * ctr1) + ctr2 * ctr2)
        proj.append(dict(c=c, s=s, SW=SW, kin=kin, units=units, ctr1=ctr1, ctr2=ctr2, cu1=ctr1 / nc, cu2=ctr2 / nc,
                         L2=(Lsd - SW[0]) / float(_cos(alpha)[0]), n=dety.size))
    out = np.zeros((len(q_rows), 4))
    matched = []
    for u, q in enumerate(np.asarray(q_rows, float).reshape(-1, 4)):
        if not ((q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]) <= 1.0:
            matched.append(np.zeros(0))
            continue
        SUB = _mm(_mm(S, _U(q)), B)
        Gw = np.array([[(SUB[r, 0] * hkl[h, 0] + SUB[r, 1] * hkl[h, 1]) + SUB[r, 2] * hkl[h, 2] for r in range(3)]
                       for h in range(H)])
        angles = []
        cost = [0.0, 0.0, 0.0]
        expect = 0
        for i, P in enumerate(proj):
            c, s = P["c"], P["s"]
            Gt = np.stack([c * Gw[:, 0] - s * Gw[:, 1], s * Gw[:, 0] + c * Gw[:, 1], Gw[:, 2]], 1)
            glen = np.sqrt((Gt[:, 0] * Gt[:, 0] + Gt[:, 1] * Gt[:, 1]) + Gt[:, 2] * Gt[:, 2])
            expect += _expect(geo, P, Gt, glen, beta_max)
            for h in range(H):
                k = fam[h]
                if k < 0 or P["n"] == 0:
                    continue
                nv = [Gt[h, 0] / glen[h], Gt[h, 1] / glen[h], Gt[h, 2] / glen[h]]
                un = P["units"][k]
                dots = (nv[0] * un[0] + nv[1] * un[1]) + nv[2] * un[2]
                ang = (np.where(dots > 1.0, 0.0, _acos(np.minimum(dots, 1.0))) * 180.0) / PI
                ang = np.where(np.isnan(dots), np.nan, ang)
                if np.all(np.isnan(ang)):
                    continue
                j = int(np.nanargmin(ang))  # first index of the minimum
                amin = float(ang[j])
                if amin < 1.0 and abs(glen[h] - glen_possible[k]) < 0.1:
                    angles.append(amin)
                    for r in range(3):
                        e = nv[r] - float(un[r][j])
                        cost[r] = cost[r] + e * e
        n = len(angles)
        if n:
            srt = sorted(angles)
            med = srt[n // 2] if n % 2 else srt[n // 2 - 1] + (srt[n // 2] - srt[n // 2 - 1]) / 2.0
            acc = 0.0
            for a_ in angles:
                acc = acc + a_
            mean = acc / n
        else:
            med = mean = math.nan
        ratio = n / expect if expect else math.nan
        matched.append(np.array(angles))
        if semantics == "match":
            out[u] = [n, ((cost[0] + cost[1]) + cost[2]) / n if n else 100.0, ratio, mean]
        else:
            out[u] = [n, med, ratio, mean]
    return (out, matched) if details else out


def _expect(geo, P, Gt, glen, beta_max):
    """calcExpectSpotNr.m:6-47 for all reflections of one projection."""
    kin, SW = P["kin"], P["SW"]
    R = np.asarray(geo.RotDet, float)
    cb = (kin[0] * (Gt[:, 0] / glen) + kin[1] * (Gt[:, 1] / glen)) + kin[2] * (Gt[:, 2] / glen)
    beta = _acos(cb)
    theta = beta - PI / 2.0
    d = ((1.0 / glen) * 2.0) * PI
    lam = (2.0 * d) * _sin(theta)
    sintth = _sin(2.0 * theta)
    nv = np.sqrt(Gt[:, 1] * Gt[:, 1] + Gt[:, 2] * Gt[:, 2])
    phix = _acos((Gt[:, 1] / nv) * P["cu1"] + (Gt[:, 2] / nv) * P["cu2"])
    phiy = phix - 2.0 * theta
    diffvec = (P["L2"] * sintth) / _sin(phiy)
    y22 = P["ctr1"] + ((diffvec * Gt[:, 1]) / nv)
    z22 = P["ctr2"] + ((diffvec * Gt[:, 2]) / nv)
    k0 = geo.Lsam2det - SW[0]
    k1, k2 = y22 - SW[1], z22 - SW[2]
    nk = np.sqrt((k0 * k0 + k1 * k1) + k2 * k2)
    K0, K1, K2 = k0 / nk, k1 / nk, k2 / nk
    tnum = (R[0, 0] * (geo.Lsam2det - SW[0]) + R[1, 0] * ((geo.dety00 - geo.RotAxisOffset) - SW[1])) + \
        R[2, 0] * (geo.detz00 - SW[2])
    t = tnum / ((R[0, 0] * K0 + R[1, 0] * K1) + R[2, 0] * K2)
    ov = [-k0, (SW[1] - geo.dety00) + geo.RotAxisOffset, SW[2] - geo.detz00]
    w0, w1, w2 = t * K0 + ov[0], t * K1 + ov[1], t * K2 + ov[2]
    Y22 = (R[0, 1] * w0 + R[1, 1] * w1) + R[2, 1] * w2
    Z22 = (R[0, 2] * w0 + R[1, 2] * w1) + R[2, 2] * w2
    dety = _mround((-Y22) / geo.pixelysize + geo.dety0)
    detz = _mround((-Z22) / geo.pixelzsize + geo.detz0)
    bs_y, bs_z = geo.BeamStopY, geo.BeamStopZ
    sel = (beta > PI / 2.0) & (beta < beta_max) & (lam > geo.lambda_min) & (lam < geo.lambda_max) & \
        (dety >= 1) & (dety <= geo.detysize) & (detz >= 1) & (detz <= geo.detzsize) & \
        ~((dety >= bs_y[0]) & (dety <= bs_y[1]) & (detz >= bs_z[0]) & (detz <= bs_z[1]))
    return int(np.count_nonzero(sel))
