"""Per-run arrays handed to the device: CPML grading, kappa axes, material maps.

* ``cpml_grading``  -- SE/KE/AE/BE/CE and SH/.../CH of
  FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/fdtd_main.py:119-168
  (FDTD_3D fdtd_main.py:128-205), evaluated on arrays.  Element types follow
  numpy 1.20: float32 op float32 stays float32, anything touching a Python
  scalar or float64 is float64, the result is stored in the run precision.
* ``kappa_axes``    -- KXE/KXH/... of fdtd_main.py:151-178 (3D :190-221).
* ``fg_mask``       -- generate_fg_mask / apply_mask (fdtd_funcs.py:165-210,
  3D :321-358) given the two roughness profiles.
* ``cb_maps``       -- EPS_MASK, average_materials (TE/3D) and CB = dt/eps
  (fdtd_main.py:83-86; TE :83-89; 3D :86-94).
"""
import numpy as np


def cpml_grading(N, m_grade, m_alpha, kappa_max, alpha_max, sigma_factor, eta, delta_x, delta_t, eps, P):
    sigma_max = sigma_factor * (0.8 * (m_grade + 1) / (float(eta) * float(delta_x)))
    d = np.arange(N, dtype=np.float64)
    g = {}
    for tag, pos, apos in (('E', (d + 0.5) / N, (N - (d + 0.5)) / N), ('H', d / N, (N - d) / N)):
        S = (sigma_max * pos ** m_grade).astype(P)
        K = (1 + (kappa_max - 1) * pos ** m_grade).astype(P)
        A = (alpha_max * apos ** m_alpha).astype(P)
        ratio = (S / K + A).astype(P)                        # float32 element arithmetic when P is f32
        B = np.exp(-1.0 * ratio.astype(np.float64) * (float(delta_t) / eps)).astype(P)
        S64, K64, A64 = S.astype(np.float64), K.astype(np.float64), A.astype(np.float64)
        den = (S * K).astype(np.float64) + K64 ** 2 * A64
        with np.errstate(invalid='ignore', divide='ignore'):
            Cc = np.where(den != 0, S64 / np.where(den != 0, den, 1.0), 0.0) * (B.astype(np.float64) - 1.0)
        g['S' + tag], g['K' + tag], g['A' + tag], g['B' + tag], g['C' + tag] = S, K, A, B, Cc.astype(P)
    return g


def kappa_axes(n, N, offset, KE, KH, delta_x, P):
    """(KE*dx, KH*dx) along one axis of length n; lo graded at [offset, offset+N]."""
    dxp = P(delta_x)
    ke = np.full(n, dxp, dtype=P)
    kh = np.full(n, dxp, dtype=P)
    e = n - offset
    d = np.arange(N)
    for dd in range(N):   # same write order as the reference loop (matters only on tiny grids)
        ke[offset + N - (dd + 1)] = KE[dd] * dxp
        kh[offset + N - dd] = KH[dd] * dxp
        ke[e - (dd + 1)] = KE[N - 1 - dd] * dxp
        kh[e - (dd + 1)] = KH[N - 1 - dd] * dxp
    del d
    return ke, kh


def padded(profile, before, after):
    return np.concatenate([np.zeros(before), np.asarray(profile, dtype=np.float64), np.zeros(after)])


def fg_mask(s, upper=None, lower=None):
    """Core mask (bool) of setup ``s`` for profile arrays of length p2-p1 (None = smooth).

    2D: core iff by + l*pp + lower[i] <= j < by + l*pp + ny_swg + upper[i].
    3D: core iff j in [by, by+ny_swg) and bx + l*pp + lower[k] <= i < bx + l*pp + nx_swg + upper[k].
    Comparisons are float64, as in apply_mask."""
    if s.scheme in ('tm', 'te'):
        L = s.p2x - s.p1x
        up = padded(np.zeros(L) if upper is None else upper, s.p1x, s.nx - s.p2x)
        lo = padded(np.zeros(L) if lower is None else lower, s.p1x, s.nx - s.p2x)
        j = np.arange(s.ny, dtype=np.float64)[None, :]
        m = np.zeros((s.nx, s.ny), dtype=bool)
        for l in range(s.num_lines):
            off = l * s.pp
            m |= (j >= s.by + off + lo[:, None]) & (j < s.by + off + s.ny_swg + up[:, None])
        return m
    L = s.p2z - s.p1z
    up = padded(np.zeros(L) if upper is None else upper, s.p1z, s.nz - s.p2z)
    lo = padded(np.zeros(L) if lower is None else lower, s.p1z, s.nz - s.p2z)
    i = np.arange(s.nx, dtype=np.float64)[:, None]
    xz = np.zeros((s.nx, s.nz), dtype=bool)
    for l in range(s.num_lines):
        off = l * s.pp
        xz |= (i >= s.bx + off + lo[None, :]) & (i < s.bx + off + s.nx_swg + up[None, :])
    jm = np.zeros(s.ny, dtype=bool)
    jm[max(s.by, 0):max(s.by + s.ny_swg, 0)] = True
    return xz[:, None, :] & jm[None, :, None]


def cb_maps(s, fg):
    """CB maps (run precision) for the E components of the scheme."""
    P = s.precision
    eps_map = s.eps * np.where(fg, s.eps_rel_fg, 1.0)
    dt = P(s.delta_t)
    if s.scheme == 'tm':
        return (dt / eps_map.astype(P),)
    avg = [eps_map.astype(P) for _ in range(2 if s.scheme == 'te' else 3)]
    inner = tuple(slice(1, None) for _ in range(eps_map.ndim))
    for a, arr in enumerate(avg):
        prev = [slice(1, None)] * eps_map.ndim
        prev[a] = slice(0, -1)
        arr[inner] = 0.5 * (eps_map[inner] + eps_map[tuple(prev)])
    return tuple(dt / arr for arr in avg)
