"""Oracle restatement of the reference geometry / material maps (test infrastructure only).

Follows FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/fdtd_funcs.py:165-210
(``generate_fg_mask`` with mode='load' or 'smooth', ``apply_mask``),
FDTD/GPU/FDTD_2D_TE/s-parameter_extraction/fdtd_funcs.py:147-152 and
FDTD/GPU/FDTD_3D/s-parameter_extraction/fdtd_funcs.py:295-302,345-358
(``average_materials``, 3D ``apply_mask``), and the ε/CB construction of
FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/fdtd_main.py:83-86,
FDTD/GPU/FDTD_2D_TE/s-parameter_extraction/fdtd_main.py:83-89,
FDTD/GPU/FDTD_3D/s-parameter_extraction/fdtd_main.py:86-94.
"""
import numpy as np


def padded_profiles(upper, lower, settling_range, end_range):
    # fdtd_funcs.py:194-197
    up = np.append(np.append(np.zeros(settling_range), upper), np.zeros(end_range))
    lo = np.append(np.append(np.zeros(settling_range), lower), np.zeros(end_range))
    return up, lo


def apply_mask_2d(num_x, num_y, bord_y, wg_height, pitch, num_lines, lower, upper):
    """Literal per-cell loop of apply_mask (fdtd_funcs.py:203-210). Fine for tiny grids."""
    m = np.zeros([num_x, num_y], dtype=bool)
    for l in range(num_lines):
        offset = l * pitch
        for j in range(num_y):
            for i in range(num_x):
                if j >= (bord_y + offset + lower[i]) and j < (bord_y + offset + wg_height + upper[i]):
                    m[i, j] = True
    return m


def apply_mask_2d_np(num_x, num_y, bord_y, wg_height, pitch, num_lines, lower, upper):
    """Same predicate evaluated with broadcasting (float64 compares)."""
    m = np.zeros([num_x, num_y], dtype=bool)
    j = np.arange(num_y, dtype=np.float64)[None, :]
    for l in range(num_lines):
        offset = l * pitch
        m |= (j >= (bord_y + offset + lower[:, None])) & (j < (bord_y + offset + wg_height + upper[:, None]))
    return m


def apply_mask_3d_np(num_x, num_y, num_z, bord_x, bord_y, wg_height, wg_width, pitch, num_lines, lower, upper):
    """3D apply_mask (FDTD_3D fdtd_funcs.py:345-353): the rough test is on x,
    indexed by the propagation coordinate k."""
    m = np.zeros([num_x, num_y, num_z], dtype=bool)
    i = np.arange(num_x, dtype=np.float64)[:, None, None]
    j = np.arange(num_y)[None, :, None]
    jm = (j >= bord_y) & (j < bord_y + wg_width)
    for l in range(num_lines):
        offset = l * pitch
        m |= jm & (i >= (bord_x + offset + lower[None, None, :])) & (i < (bord_x + offset + wg_height + upper[None, None, :]))
    return m


def fg_mask(c, upper=None, lower=None, line_count=None):
    """FG_REG for setup namespace ``c``; smooth when profiles are None."""
    L = c.num_lines if line_count is None else line_count
    if c.scheme in ('tm', 'te'):
        neff = c.p2x - c.p1x
        upper = np.zeros(neff) if upper is None else np.asarray(upper, np.float64)
        lower = np.zeros(neff) if lower is None else np.asarray(lower, np.float64)
        up, lo = padded_profiles(upper, lower, c.p1x, c.nx - c.p2x)
        return apply_mask_2d_np(c.nx, c.ny, c.by, c.ny_swg, c.pp, L, lo, up)
    neff = c.p2z - c.p1z
    upper = np.zeros(neff) if upper is None else np.asarray(upper, np.float64)
    lower = np.zeros(neff) if lower is None else np.asarray(lower, np.float64)
    up, lo = padded_profiles(upper, lower, c.p1z, c.nz - c.p2z)
    return apply_mask_3d_np(c.nx, c.ny, c.nz, c.bx, c.by, c.nx_swg, c.ny_swg, c.pp, L, lo, up)


def eps_mask(c, FG):
    # fdtd_main.py:83 -- float64 permittivity map
    return c.eps * (c.eps_rel_fg * FG + 1 * ~FG)


def average_materials_2d(Mask, X, Y, ie, je):
    """fdtd_funcs.py (TE):147-152, vectorised; X, Y modified in place."""
    X[1:ie, 1:je] = 0.5 * (Mask[1:ie, 1:je] + Mask[0:ie - 1, 1:je])
    Y[1:ie, 1:je] = 0.5 * (Mask[1:ie, 1:je] + Mask[1:ie, 0:je - 1])


def average_materials_3d(Mask, X, Y, Z, ie, je, ke):
    """fdtd_funcs.py (3D):295-302, vectorised; X, Y, Z modified in place."""
    s = (slice(1, ie), slice(1, je), slice(1, ke))
    X[s] = 0.5 * (Mask[s] + Mask[0:ie - 1, 1:je, 1:ke])
    Y[s] = 0.5 * (Mask[s] + Mask[1:ie, 0:je - 1, 1:ke])
    Z[s] = 0.5 * (Mask[s] + Mask[1:ie, 1:je, 0:ke - 1])


def cb_maps(c, FG):
    """CB arrays in the run precision: TM -> (CBZ,), TE -> (CBX, CBY), 3D -> (CBX, CBY, CBZ)."""
    P = c.precision
    M = eps_mask(c, FG)
    dt = P(c.delta_t)
    if c.scheme == 'tm':
        return (dt / M.astype(P),)
    if c.scheme == 'te':
        X = M.astype(P); Y = M.astype(P)
        average_materials_2d(M, X, Y, c.nx, c.ny)
        return (dt / X, dt / Y)
    X = M.astype(P); Y = M.astype(P); Z = M.astype(P)
    average_materials_3d(M, X, Y, Z, c.nx, c.ny, c.nz)
    return (dt / X, dt / Y, dt / Z)
