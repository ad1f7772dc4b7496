"""Oracle restatement of the reference S-parameter extraction (test infrastructure only).

Follows FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/extract_sparams.py:50-104
(2D TM; TE uses +0.5 Ey Hz*, FDTD_2D_TE/.../extract_sparams.py:88-89) and
FDTD/GPU/FDTD_3D/s-parameter_extraction/extract_sparams.py:50-131 (3D), with
the accumulation loops written out as in the reference and complex128 math.
"""
import numpy as np


def _sav(scheme, f):
    if scheme == 'tm':
        return -0.5 * (f[0] * np.conjugate(f[1]))
    if scheme == 'te':
        return 0.5 * (f[0] * np.conjugate(f[1]))
    ex, ey, hx, hy = f
    return 0.5 * (ex * np.conjugate(hy) - ey * np.conjugate(hx))


def s_matrix(scheme, c, files, num_lines, nf):
    """``files[(tag, line)]`` -> tuple of component phasors (TM: ez, hy; TE: ey,
    hz; 3D: ex, ey, hx, hy).  ``c`` carries by/bx, cy_swg/cx_swg, ph, pp, ny,
    num_cpml."""
    P = 2 * num_lines
    S = np.zeros([nf, P, P], dtype=np.complex128)
    for i in range(P):
        for j in (range(P) if scheme == '3d' else range(i, P)):
            iline, jline = i // 2 + 1, j // 2 + 1
            inc = files[('fi' if j % 2 == 0 else 'bi', jline)]
            if i % 2 == j % 2:
                ref = files[('ft' if i % 2 == 0 else 'bt', jline)]
            else:
                ref = files[('fr' if i % 2 == 0 else 'br', jline)]
            ref = tuple(np.array(a, dtype=np.complex128) for a in ref)
            if i == j:
                cor = files[('fi' if i % 2 == 0 else 'bi', iline)]
                ref = tuple(a - b for a, b in zip(ref, cor))
            sav_inc = _sav(scheme, inc)
            sav_ref = _sav(scheme, ref)
            if scheme == '3d':
                start_inc = c.bx + c.cx_swg - c.ph + c.pp * (jline - 1)
                start_ref = c.bx + c.cx_swg - c.ph + c.pp * (iline - 1)
            else:
                start_inc = c.by + c.cy_swg - c.ph + c.pp * (jline - 1)
                start_ref = c.by + c.cy_swg - c.ph + c.pp * (iline - 1)
            Pinc = np.zeros(nf, dtype=np.complex128)
            Pref = np.zeros(nf, dtype=np.complex128)
            if scheme == '3d':
                for y in range(c.num_cpml, c.ny - c.num_cpml):
                    for p in range(start_inc, start_inc + c.pp):
                        Pinc += sav_inc[:, p, y]
                    for p in range(start_ref, start_ref + c.pp):
                        Pref += sav_ref[:, p, y]
            else:
                for p in range(start_inc, start_inc + c.pp):
                    Pinc += sav_inc[:, p]
                for p in range(start_ref, start_ref + c.pp):
                    Pref += sav_ref[:, p]
            S[:, i, j] = np.sqrt(Pref / Pinc)
            S[:, j, i] = np.sqrt(Pref / Pinc)
    return S
