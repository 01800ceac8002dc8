"""Per-slab temporal data of the space-time solve (host, numpy, vectorised over load cases).

* ``stcoeff_m`` -- closed-form 6x6 temporal matrices of one slab (stcoeff.f90:62-356):
  slab matrices k, m, jump matrices at t_n(+) ktn, mtn and at t_n(-) ktn1, mtn1.
  The operator of the slab solve is A = k + ktn, B = m + mtn (main3d.f90:831) and the
  jump term uses (ktn1, mtn1) (main3d.f90:818).  The reference's default-real
  literals (2./3., 1./6., ...) are FP32 values widened to FP64; they are reproduced
  through ``_f32``.  Every argument may be an array: the result has the broadcast
  shape + (6, 6), one set of matrices per load case (frequency, step).
* ``stcoeff_t`` -- 3x3 thermal slab matrices (stcoeff.f90:1-60).
* ``getfextc_m`` -- temporal load coefficients fextc, fint (getfextc.f90:22-58).
* ``damage_nt`` -- temporal interpolation table of the damage update
  (kokkos_damage.f90:36-47).
* ``slab_rhs`` -- the right-hand side of one mechanical slab (main3d.f90:803-823).
* ``getfextc_t`` -- temporal coefficients of the heat flux (getfextc.f90:1-20).
* ``thermal_rhs`` -- the right-hand side of one thermal slab (main3d.f90:703-723).
"""
from __future__ import annotations

import numpy as np


def _f32(num, den):
    """A Fortran default-real constant num./den. (rounded to FP32, then widened)."""
    return float(np.float32(np.float32(num) / np.float32(den)))


def stcoeff_m(t1, dt, w):
    t1, dt, w = np.broadcast_arrays(np.asarray(t1, dtype=np.float64),
                                    np.asarray(dt, dtype=np.float64),
                                    np.asarray(w, dtype=np.float64))
    sh = t1.shape
    sin, cos = np.sin, np.cos
    dt2 = dt * dt
    dt3 = dt2 * dt
    dt4 = dt2 * dt2
    w2 = w * w
    w3 = w2 * w
    w4 = w2 * w2
    a1 = t1 * w                      # phase at the slab start
    a2 = (dt + t1) * w               # phase at the slab end
    ah = (dt / 2. + t1) * w          # phase at the slab midpoint
    S1, C1, S2, C2, Sh = sin(a1), cos(a1), sin(a2), cos(a2), sin(ah)
    S2t, S2e = sin(2. * t1 * w), sin(2. * (dt + t1) * w)
    C2t, C2e = cos(2. * t1 * w), cos(2. * (dt + t1) * w)
    Cd, Sd, Cm = cos(dt * w), sin(dt * w), cos((dt + 2. * t1) * w)
    dw = dt * w
    q4 = dt4 * w4
    p2 = dt2 * w2

    k = np.zeros(sh + (6, 6))
    k[..., 0, 0] = -0.5
    k[..., 0, 1] = -_f32(2, 3)
    k[..., 0, 2] = _f32(1, 6)
    k[..., 0, 3] = (-6. * dw * (-12. + p2) * C1 + 24. * dw * C2 + (96. - 26. * p2 + q4) * S1
                    + 2. * (-48. + p2) * S2) / (2. * q4)
    k[..., 0, 4] = (2. * (dw * (-84. * C1 - 60. * C2 + dw * (18. * S1 + p2 * Sh - 6. * S2))
                          - 144. * (S1 - S2))) / (3. * q4)
    k[..., 0, 5] = (120. * dw * C1 - 6. * dw * (-28. + p2) * C2 - 18. * (-16. + p2) * S1
                    - (288. - 42. * p2 + q4) * S2) / (6. * q4)
    k[..., 1, 0] = _f32(2, 3)
    k[..., 1, 2] = -_f32(2, 3)
    k[..., 1, 3] = (-2. * (-6. * dw * (-16. + p2) * C1 + 48. * dw * C2
                           + (-24. + p2) * ((-6. + p2) * S1 + 6. * S2))) / (3. * q4)
    k[..., 1, 4] = (16. * (6. * dw * C1 + 6. * dw * C2 - (-12. + p2) * (S1 - S2))) / q4
    k[..., 1, 5] = (2. * (-48. * dw * C1 + 6. * dw * (-16. + p2) * C2
                          + (-24. + p2) * (6. * S1 + (-6. + p2) * S2))) / (3. * q4)
    k[..., 2, 0] = -_f32(1, 6)
    k[..., 2, 1] = _f32(2, 3)
    k[..., 2, 2] = 0.5
    k[..., 2, 3] = (-6. * dw * (-28. + p2) * C1 + 120. * dw * C2 + (288. - 42. * p2 + q4) * S1
                    + 18. * (-16. + p2) * S2) / (6. * q4)
    k[..., 2, 4] = (-2. * (144. * S1 + dw * (60. * C1 + 84. * C2 - 6. * dw * S1 + dt3 * w3 * Sh)
                           + 18. * (-8. + p2) * S2)) / (3. * q4)
    k[..., 2, 5] = (96. * (S1 - S2) + dw * (24. * C1 + (72. - 6. * p2) * C2
                                            + dw * (-2. * S1 + (26. - p2) * S2))) / (2. * q4)
    k[..., 3, 4] = (576. - 72. * p2 - 4. * q4 + 6. * (-96. + 13. * p2) * C1 * C1
                    + 576. * S1 * S1 - 78. * p2 * S1 * S1 + 339. * dw * S2t
                    - 6. * dt3 * w3 * S2t + 96. * dw * C2 * (5. * S1 - 4. * Sh)
                    + 48. * dw * (-16. + p2) * C1 * Sh - 1152. * S1 * Sh + 240. * p2 * S1 * Sh
                    - 8. * q4 * S1 * Sh - 1152. * S1 * S2 + 48. * p2 * S1 * S2 + 1152. * Sh * S2
                    - 48. * p2 * Sh * S2 - 3. * dw * S2e) / (12. * q4)
    k[..., 3, 5] = (-1152. + 72. * p2 + 2. * q4 + 2. * (576. - 84. * p2 + q4) * Cd
                    - 6. * (-96. + 7. * p2) * C2t + 576. * C2e - 42. * p2 * C2e - 1152. * Cm
                    + 168. * p2 * Cm - 2. * q4 * Cm + 672. * dw * Sd - 24. * dt3 * w3 * Sd
                    - 246. * dw * S2t + 3. * dt3 * w3 * S2t + 246. * dw * S2e
                    - 3. * dt3 * w3 * S2e) / (24. * q4)
    k[..., 4, 5] = -(-576. + 72. * p2 + 4. * q4 + (576. - 78. * p2) * C2 * C2 - 3. * dw * S2t
                     - 384. * dw * C1 * Sh + 48. * dw * (-16. + p2) * C2 * Sh - 1152. * S1 * Sh
                     + 48. * p2 * S1 * Sh + 480. * dw * C1 * S2 + 1152. * S1 * S2
                     - 48. * p2 * S1 * S2 + 1152. * Sh * S2 - 240. * p2 * Sh * S2
                     + 8. * q4 * Sh * S2 - 576. * S2 * S2 + 78. * p2 * S2 * S2 + 339. * dw * S2e
                     - 6. * dt3 * w3 * S2e) / (12. * q4)
    # rows 4..6 of k are minus the transposed columns 4..6 (stcoeff.f90:124-177)
    for i in range(3, 6):
        for j in range(0, i):
            k[..., i, j] = -k[..., j, i]

    m = np.zeros(sh + (6, 6))
    m[..., 0, 0] = -4. / dt2
    m[..., 0, 1] = 8. / dt2
    m[..., 0, 2] = -4. / dt2
    m[..., 0, 3] = (3. * dw * C1 - S1 + S2) / dt2
    m[..., 0, 4] = (4. * (3. * S1 - 2. * Sh - S2)) / dt2
    m[..., 0, 5] = (dw * C2 - 3. * S1 + 3. * S2) / dt2
    m[..., 1, 3] = (-4. * (dw * C1 - S1 + S2)) / dt2
    m[..., 1, 4] = (16. * (-S1 + S2)) / dt2
    m[..., 1, 5] = (-4. * (dw * C2 - S1 + S2)) / dt2
    m[..., 2, 0] = 4. / dt2
    m[..., 2, 1] = -8. / dt2
    m[..., 2, 2] = 4. / dt2
    m[..., 2, 3] = (dw * C1 - 3. * S1 + 3. * S2) / dt2
    m[..., 2, 4] = (4. * (S1 + 2. * Sh - 3. * S2)) / dt2
    m[..., 2, 5] = (3. * dw * C2 - S1 + S2) / dt2
    m[..., 3, 3] = -(-2. + p2 + 2. * Cd + (1. + p2) * C2t + C2e - 2. * Cm) / (4. * dt2)
    m[..., 3, 4] = (-8. * dw + 4. * dt3 * w3 + 3. * S2t
                    + 2. * dw * (4. * Cd + 5. * C2t + 2. * C2e - 4. * Cm - 3. * dw * S2t)
                    - 3. * S2e) / (4. * dt3 * w)
    m[..., 3, 5] = (-16. * dw - 2. * dt3 * w3 + 8. * dw * C1 * C1 - 4. * dw * C2 * C2
                    - 8. * p2 * C2 * S1 + 16. * dw * S1 * S1 - 6. * S2t + 3. * p2 * S2t
                    + 8. * dw * S1 * S2 + 4. * dw * S2 * S2 + 6. * S2e + p2 * S2e) / (8. * dt3 * w)
    m[..., 4, 3] = -(4. * dt3 * w3 + 10. * dw * C1 * C1 - 4. * dw * C2 * C2 - 10. * dw * S1 * S1
                     + 3. * S2t + 2. * p2 * S2t - 16. * p2 * C1 * Sh + 16. * dw * S1 * Sh
                     - 16. * dw * Sh * S2 + 4. * dw * S2 * S2 - 3. * S2e) / (4. * dt3 * w)
    m[..., 4, 4] = (4. * (C2t - C2e + 2. * cos(((3 * dt) / 2. + 2. * t1) * w)
                          - 2. * cos(((dt + 4. * t1) * w) / 2.))) / dt2
    m[..., 4, 5] = (4. * dt3 * w3 - 4. * dw * C1 * C1 + 10. * dw * C2 * C2 + 4. * dw * S1 * S1
                    + 3. * S2t + 16. * p2 * C2 * Sh - 16. * dw * S1 * Sh + 16. * dw * Sh * S2
                    - 10. * dw * S2 * S2 - 3. * S2e - 2. * p2 * S2e) / (4. * dt3 * w)
    m[..., 5, 3] = (2. * dw + 2. * dt3 * w3 - 4. * dw * Cd + 4. * dw * C1 * C1 + 2. * dw * C2t
                    + 4. * dw * C2e + 4. * dw * Cm + 6. * S2t + p2 * S2t - 8. * p2 * C1 * S2
                    - 6. * S2e + 3. * p2 * S2e) / (8. * dt3 * w)
    m[..., 5, 4] = (8. * dw - 4. * dt3 * w3 - 3. * S2t + 3. * S2e
                    - 2. * dw * (4. * Cd + 2. * C2t + 5. * C2e - 4. * Cm + 3. * dw * S2e)) / (4. * dt3 * w)
    m[..., 5, 5] = (-2. + p2 + 2. * Cd + C2t + (1. + p2) * C2e - 2. * Cm) / (4. * dt2)

    ktn = np.zeros(sh + (6, 6))
    ktn[..., 0, 0] = 1.
    # jump mass matrix at t_n(+) is symmetric (stcoeff.f90:264-304)
    mtn = np.zeros(sh + (6, 6))
    mtn[..., 0, 0] = 9. / dt2
    mtn[..., 0, 1] = -12. / dt2
    mtn[..., 0, 2] = 3. / dt2
    mtn[..., 0, 3] = (-3. * w * C1) / dt
    mtn[..., 0, 4] = (12. * (-S1 + Sh)) / dt2
    mtn[..., 0, 5] = (3. * (S1 - S2)) / dt2
    mtn[..., 1, 1] = 16. / dt2
    mtn[..., 1, 2] = -4. / dt2
    mtn[..., 1, 3] = (4. * w * C1) / dt
    mtn[..., 1, 4] = (16. * (S1 - Sh)) / dt2
    mtn[..., 1, 5] = (4. * (-S1 + S2)) / dt2
    mtn[..., 2, 2] = 1. / dt2
    mtn[..., 2, 3] = -((w * C1) / dt)
    mtn[..., 2, 4] = (4. * (-S1 + Sh)) / dt2
    mtn[..., 2, 5] = (S1 - S2) / dt2
    mtn[..., 3, 3] = w2 * C1 * C1
    mtn[..., 3, 4] = (4. * w * C1 * (S1 - Sh)) / dt
    mtn[..., 3, 5] = (w * C1 * (-S1 + S2)) / dt
    mtn[..., 4, 4] = (16. * (S1 - Sh) * (S1 - Sh)) / dt2
    mtn[..., 4, 5] = (4. * (S1 - Sh) * (-S1 + S2)) / dt2
    mtn[..., 5, 5] = (S1 - S2) * (S1 - S2) / dt2
    for i in range(6):
        for j in range(i):
            mtn[..., i, j] = mtn[..., j, i]

    ktn1 = np.zeros(sh + (6, 6))
    ktn1[..., 0, 2] = 1.
    mtn1 = np.zeros(sh + (6, 6))
    mtn1[..., 0, 0] = -3. / dt2
    mtn1[..., 0, 1] = 12. / dt2
    mtn1[..., 0, 2] = -9. / dt2
    mtn1[..., 0, 3] = (-3. * (-S1 + S2)) / dt2
    mtn1[..., 0, 4] = (-12. * (Sh - S2)) / dt2
    mtn1[..., 0, 5] = (-3. * w * C2) / dt
    mtn1[..., 1, 0] = 4. / dt2
    mtn1[..., 1, 1] = -16. / dt2
    mtn1[..., 1, 2] = 12. / dt2
    mtn1[..., 1, 3] = (4. * (-S1 + S2)) / dt2
    mtn1[..., 1, 4] = (16. * (Sh - S2)) / dt2
    mtn1[..., 1, 5] = (4. * w * C2) / dt
    mtn1[..., 2, 0] = -1. / dt2
    mtn1[..., 2, 1] = 4. / dt2
    mtn1[..., 2, 2] = -3. / dt2
    mtn1[..., 2, 3] = -((-S1 + S2) / dt2)
    mtn1[..., 2, 4] = (-4. * (Sh - S2)) / dt2
    mtn1[..., 2, 5] = -((w * C2) / dt)
    mtn1[..., 3, 0] = (w * C1) / dt
    mtn1[..., 3, 1] = (-4. * w * C1) / dt
    mtn1[..., 3, 2] = (3. * w * C1) / dt
    mtn1[..., 3, 3] = (w * C1 * (-S1 + S2)) / dt
    mtn1[..., 3, 4] = (4. * w * C1 * (Sh - S2)) / dt
    mtn1[..., 3, 5] = w2 * C1 * C2
    mtn1[..., 4, 0] = (4. * (S1 - Sh)) / dt2
    mtn1[..., 4, 1] = (-16. * (S1 - Sh)) / dt2
    mtn1[..., 4, 2] = (12. * (S1 - Sh)) / dt2
    mtn1[..., 4, 3] = (4. * (S1 - Sh) * (-S1 + S2)) / dt2
    mtn1[..., 4, 4] = (16. * (S1 - Sh) * (Sh - S2)) / dt2
    mtn1[..., 4, 5] = (4. * w * C2 * (S1 - Sh)) / dt
    mtn1[..., 5, 0] = (-S1 + S2) / dt2
    mtn1[..., 5, 1] = (-4. * (-S1 + S2)) / dt2
    mtn1[..., 5, 2] = (3. * (-S1 + S2)) / dt2
    mtn1[..., 5, 3] = (-S1 + S2) * (-S1 + S2) / dt2
    mtn1[..., 5, 4] = (4. * (Sh - S2) * (-S1 + S2)) / dt2
    mtn1[..., 5, 5] = (w * C2 * (-S1 + S2)) / dt
    return {"k": k, "m": m, "ktn": ktn, "mtn": mtn, "ktn1": ktn1, "mtn1": mtn1}


def stcoeff_t(t1, dt):
    """3x3 thermal slab matrices (stcoeff.f90:1-60); t1 is unused by the reference."""
    dt = np.asarray(dt, dtype=np.float64)
    sh = dt.shape
    k = np.zeros(sh + (3, 3))
    vals = [[(2, 15), (1, 15), (-1, 30)], [(1, 15), (8, 15), (1, 15)], [(-1, 30), (1, 15), (2, 15)]]
    for i in range(3):
        for j in range(3):
            nm, dn = vals[i][j]
            c = _f32(abs(nm), dn)
            k[..., i, j] = (c if nm > 0 else -c) * dt
    m = np.zeros(sh + (3, 3))
    m[..., 0, :] = [-0.5, _f32(2, 3), -_f32(1, 6)]
    m[..., 1, :] = [-_f32(2, 3), 0.0, _f32(2, 3)]
    m[..., 2, :] = [_f32(1, 6), -_f32(2, 3), 0.5]
    ktn = np.zeros(sh + (3, 3))
    mtn = np.zeros(sh + (3, 3))
    mtn[..., 0, 0] = 1.
    ktn1 = np.zeros(sh + (3, 3))
    mtn1 = np.zeros(sh + (3, 3))
    mtn1[..., 0, 2] = 1.
    return {"k": k, "m": m, "ktn": ktn, "mtn": mtn, "ktn1": ktn1, "mtn1": mtn1}


def getfextc_m(t1, w, dt):
    """Temporal load coefficients (getfextc.f90:22-58) -> fextc[..., 6], fint[..., 6]."""
    t1, w, dt = np.broadcast_arrays(np.asarray(t1, dtype=np.float64),
                                    np.asarray(w, dtype=np.float64),
                                    np.asarray(dt, dtype=np.float64))
    sin, cos = np.sin, np.cos
    dt2, w2 = dt * dt, w * w
    dw = dt * w
    S1, C1 = sin(t1 * w), cos(t1 * w)
    S2, C2 = sin((dt + t1) * w), cos((dt + t1) * w)
    f = np.zeros(t1.shape + (6,))
    f[..., 0] = -((3. * dw * C1 + dw * C2 + 4. * S1 - 4. * S2) / (dt2 * w2))
    f[..., 1] = (4. * (dw * C1 + dw * C2 + 2. * S1 - 2. * S2)) / (dt2 * w2)
    f[..., 2] = -((dw * C1 + 3. * dw * C2 + 4. * S1 - 4. * S2) / (dt2 * w2))
    f[..., 3] = (16. - 4. * dt2 * w2 + 2. * (-7. + dt2 * w2) * cos(2. * t1 * w)
                 - 2. * cos(2. * (dt + t1) * w) + 8. * dw * C2 * S1 + 9. * dw * sin(2. * t1 * w)
                 - 32. * S1 * S2 - dw * sin(2. * (dt + t1) * w)) / (8. * dt2 * w2)
    f[..., 4] = (-cos(2. * t1 * w) + cos(2. * (dt + t1) * w) + 8. * cos(((dt + 4. * t1) * w) / 2.)
                 - 8. * cos((3. * dt * w) / 2. + 2. * t1 * w)
                 + dw * (sin(2. * t1 * w) + sin(2. * (dt + t1) * w)
                         - 4. * (sin(((dt + 4. * t1) * w) / 2.) + sin((3. * dt * w) / 2. + 2. * t1 * w))))\
        / (2. * dt2 * w2)
    f[..., 5] = (-16. + 4. * dt2 * w2 + 2. * cos(2. * t1 * w)
                 - 2. * (-7. + dt2 * w2) * cos(2. * (dt + t1) * w) - dw * sin(2. * t1 * w)
                 + 8. * (dw * C1 + 4. * S1) * S2 + 9. * dw * sin(2. * (dt + t1) * w)) / (8. * dt2 * w2)
    g = np.zeros(t1.shape + (6,))
    g[..., 0] = -1.
    g[..., 2] = 1.
    g[..., 5] = -S1 + sin(w * (t1 + dt))
    return f, g


# Angular frequency of the thermal load, w = 2*3.14159/24 evaluated in default real
# (FP32) arithmetic and widened (getfextc.f90:9).
W_THERMAL = float(np.float32(np.float32(2.0 * np.float32(3.14159)) / np.float32(24.0)))


def getfextc_t(t1, dt):
    """Temporal load coefficients of the heat flux (getfextc.f90:1-20) -> fextc[..., 3],
    with the reference's operand order ((3 dt) w cos(...) etc.)."""
    t1, dt = np.broadcast_arrays(np.asarray(t1, dtype=np.float64),
                                 np.asarray(dt, dtype=np.float64))
    w = W_THERMAL
    C1, C2 = np.cos(t1 * w), np.cos((dt + t1) * w)
    S1, S2 = np.sin(t1 * w), np.sin((dt + t1) * w)
    den = (dt * dt) * (w * w)
    f = np.zeros(t1.shape + (3,))
    f[..., 0] = -((3. * dt * w * C1 + dt * w * C2 + 4. * S1 - 4. * S2) / den)
    f[..., 1] = (4. * (dt * w * C1 + dt * w * C2 + 2. * S1 - 2. * S2)) / den
    f[..., 2] = -((dt * w * C1 + 3. * dt * w * C2 + 4. * S1 - 4. * S2) / den)
    return f


def thermal_rhs(p, fextc, rext, jump=None, init=None, bound=None, fixed=None):
    """Right-hand side of one thermal slab, temporal-major [3*p] (main3d.f90:703-723):
    ``init`` (fInit, first slab only) plus block j = fextc(j) rext, plus the jump
    (ktn1 (x) K_t + mtn1 (x) M_t) T_prev, plus ``bound`` (fBound), then zero at the fixed
    space-time DOFs."""
    f = np.zeros(3 * p) if init is None else np.array(init, dtype=np.float64)
    for j in range(3):
        f[j * p:(j + 1) * p] = f[j * p:(j + 1) * p] + fextc[j] * rext
    if jump is not None:
        f = f + jump
    if bound is not None:
        f = f + bound
    if fixed is not None:
        f[np.asarray(fixed) != 0] = 0.0
    return f


def damage_nt(tcoorde, w, nip):
    """Nt[nip, 6] at t_i = t1 + (i-1) dt/nip (kokkos_damage.f90:36-47)."""
    tc = np.asarray(tcoorde, dtype=np.float64)
    dt = tc[2] - tc[0]
    t = tc[0] + np.arange(nip) * dt / nip
    N1 = 2.0 * (tc[1] - t) * (tc[2] - t) / (dt * dt)
    N2 = -4.0 * (tc[0] - t) * (tc[2] - t) / (dt * dt)
    N3 = 2.0 * (tc[0] - t) * (tc[1] - t) / (dt * dt)
    st = np.sin(w * t)
    return np.stack([N1, N2, N3, N1 * (st - np.sin(w * tc[0])), N2 * (st - np.sin(w * tc[1])),
                     N3 * (st - np.sin(w * tc[2]))], axis=1)


def slab_rhs(p, fextc, fint, rext, vv=None, jump=None, fixed=None):
    """Right-hand side of one mechanical slab, temporal-major [6*p]
    (main3d.f90:803-823): block j = fextc(j) rext + fint(j) vv, plus the jump term
    (ktn1 (x) K + mtn1 (x) M) x_prev if given, then zero at fixed space-time DOFs."""
    f = np.zeros(6 * p)
    for j in range(6):
        blk = fextc[j] * rext
        if vv is not None:
            blk = blk + fint[j] * vv
        f[j * p:(j + 1) * p] = blk
    if jump is not None:
        f = f + jump
    if fixed is not None:
        f[np.asarray(fixed) != 0] = 0.0
    return f
