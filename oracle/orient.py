"""Orientation bookkeeping after an orientation fit -- TEST INFRASTRUCTURE ONLY
(checker of the engine's grs_fit_orientations and of the indexing rows).

fit_OR_constrain_with_local_comp_cuda.m:133-148:
  [yaw,pitch,roll] = rod2angle(quat2rod([sqrt(1-x(1)^2-x(2)^2-x(3)^2) x]),'ZXZ')
  (degrees, +360 when negative); rod = angle2rod(yaw*pi/180, pitch*pi/180,
  roll*pi/180,'ZXZ'); quat = rod2quat(rod); U = quaternion2U(quat).
Restated from quat2rod.m:4-14 (quatnormalize.m, quatmod.m, quatnorm.m: the
quaternion divided by sqrt(sum(q.^2))), rod2angle.m ('zxz' case, twoaxisrot),
angle2rod.m ('zxz' case and :120-126), rod2quat.m and quaternion2U.m, every
statement in MATLAB's left-to-right order.  sin, cos, acos and atan are the
engine's FP64 functions evaluated through the oracle library; atan2 and tan
are built from them as the engine defines them (atan2 from atan of the
smaller-over-larger ratio with the quadrant fixed up, tan = sin/cos); sums of
squares and norm() are sequential.  rod2angle's gimbal-lock branch
(dcm2angle, not part of the reference) and complex intermediates give None.
"""
from __future__ import annotations

import math

import numpy as np

from . import math_eval

PIO2_HI = float.fromhex("0x1.921fb54442d18p+0")
PIO2_LO = float.fromhex("0x1.1a62633145c07p-54")
PI_HI = float.fromhex("0x1.921fb54442d18p+1")
PI_LO = float.fromhex("0x1.1a62633145c07p-53")
PI = PI_HI


def _f(fn: int, x: float) -> float:
    return float(math_eval(fn, np.array([x]))[0])


def sin(x):
    return _f(0, x)


def cos(x):
    return _f(1, x)


def acos(x):
    return _f(2, x)


def atan(x):
    return _f(3, x)


def atan2(y: float, x: float) -> float:
    if math.isnan(x) or math.isnan(y):
        return x + y
    ax, ay = abs(x), abs(y)
    if ax == 0.0 and ay == 0.0:
        r = 0.0
    elif ay <= ax:
        r = atan(ay / ax)
    else:
        r = (PIO2_HI - atan(ax / ay)) + PIO2_LO
    if x < 0.0:
        r = (PI_HI - r) + PI_LO
    return -r if y < 0.0 else r


def tan(x: float) -> float:
    return sin(x) / cos(x)


def fit_orientation(x):
    """Returns (out10 = [yaw pitch roll, rod(3), quat(4)], U (3, 3)) or None."""
    x = [float(v) for v in x]
    r0 = ((1.0 - x[0] * x[0]) - x[1] * x[1]) - x[2] * x[2]
    if not r0 >= 0.0:
        return None
    q = [math.sqrt(r0), x[0], x[1], x[2]]
    qm = math.sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3])
    qn = [v / qm for v in q]
    thalf = acos(qn[0])
    if math.isnan(thalf):
        return None
    R = [0.0, 0.0, 0.0]
    if thalf != 0.0:
        ct = cos(thalf)
        R = [qn[1] / ct, qn[2] / ct, qn[3] / ct]
    r1 = r2 = r3 = 0.0
    n = math.sqrt((R[0] * R[0] + R[1] * R[1]) + R[2] * R[2])
    if n != 0.0:
        th = 2.0 * atan(n)
        sx, sy, sz = R[0] / n, R[1] / n, R[2] / n
        cth, sth = cos(th), sin(th)
        a11 = (sx * sz) * (1.0 - cth) + sy * sth
        a12 = sx * sth - (sy * sz) * (1.0 - cth)
        a21 = sz * sz + (1.0 - sz * sz) * cth
        a31 = (sx * sz) * (1.0 - cth) - sy * sth
        a32 = (sy * sz) * (1.0 - cth) + sx * sth
        if (a11 == 0.0 and a12 == 0.0) or (a31 == 0.0 and a32 == 0.0):
            return None
        r1 = atan2(a11, a12)
        a21 = min(max(a21, -1.0), 1.0)
        r2 = acos(a21)
        r3 = atan2(a31, a32)
    ang = [(r * 180.0) / PI for r in (r1, r2, r3)]
    ang = [a + 360.0 if a < 0.0 else a for a in ang]
    rq = angle_to_rod_quat(ang)
    if rq is None:
        return None
    rod, qo = rq
    q0, q1, q2, q3 = qo
    U = np.array([
        [2.0 * (q0 * q0 + q1 * q1) - 1.0, 2.0 * (q1 * q2 - q0 * q3), 2.0 * (q1 * q3 + q0 * q2)],
        [2.0 * (q1 * q2 + q0 * q3), 2.0 * (q0 * q0 + q2 * q2) - 1.0, 2.0 * (q2 * q3 - q0 * q1)],
        [2.0 * (q1 * q3 - q0 * q2), 2.0 * (q2 * q3 + q0 * q1), 2.0 * (q0 * q0 + q3 * q3) - 1.0],
    ])
    return np.array(ang + rod + qo), U


def angle_to_rod_quat(ang):
    """rod = angle2rod(ang*pi/180, 'ZXZ') and quat = rod2quat(rod) for ZXZ
    angles in degrees (fit_OR_constrain_with_local_comp_cuda.m:139-140,
    GrainMapping_engine_gpu_comp.m:151-152).  Returns (rod (3), quat (4))
    as lists, or None where acos((trace-1)/2) would be complex."""
    b1, b2, b3 = [(float(a) * PI) / 180.0 for a in ang]
    cps, sps = cos(b1), sin(b1)
    cth, sth = cos(b2), sin(b2)
    cph, sph = cos(b3), sin(b3)
    tr = (((cph * cps - (sph * cth) * sps) - sph * sps) + (cph * cth) * cps) + cth
    e31 = sth * sps - sph * sth
    e23 = cph * sth + sth * cps
    e12 = ((cph * sps + (sph * cth) * cps) + sph * cps) + (cph * cth) * sps
    c = (tr - 1.0) / 2.0
    if not (-1.0 <= c <= 1.0):
        return None
    th = acos(c)
    rod = [0.0, 0.0, 0.0]
    if th != 0.0:
        d = 2.0 * sin(th)
        sy, sx, sz = e31 / d, e23 / d, e12 / d
        t = tan(th / 2.0)
        rod = [t * sx, t * sy, t * sz]
    qo = [0.0, 0.0, 0.0, 0.0]
    nr = math.sqrt((rod[0] * rod[0] + rod[1] * rod[1]) + rod[2] * rod[2])
    if nr != 0.0:
        th2 = atan(nr)
        s2 = sin(th2)
        qo = [cos(th2)] + [(r * s2) / nr for r in rod]
    return rod, qo
