"""Roughness ensembles with shared incident runs.

The reference extracts S-parameters per line from four runs (forward and
backward, incident and rough: FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/
simulation_launcher.py:26-35), and an ensemble study repeats the protocol for
every roughness realisation.  The incident run of a protocol sees a smooth,
shortened guide (fdtd_auto_setup.py:82-91): it depends only on the cross-
section, the port settling lengths, dt, nt and the frequencies, never on the
roughness.  ``run`` therefore simulates one incident run per distinct incident
description and pairs it with every realisation that shares it, in one device
batch with the rough runs.  S follows extract_sparams.py:52-104 (3D
FDTD_3D/.../extract_sparams.py:52-131), with the port-window power sums done
on the device (``Batch.port_power``):

* S11 = sqrt(P1(rough - incident) / P1(incident)), port 1 window;
* S21 = sqrt(P2(rough) / P1(incident)), the forward rough run's port-2
  phasors ('br' in the reference's file names) over the incident power.

Backward runs (S22, S12) follow the same pairing with ``directions=('f', 'b')``.
"""
import json

import numpy as np

from . import configs, materials
from .extract import window_start
from .solver import Batch


def incident_key(p):
    """Canonical text of the incident description of ``p``: equal keys mean
    bit-identical incident runs."""
    q = configs.incident(p)
    q.pop('seed', None)
    return json.dumps({k: (float(v) if isinstance(v, (np.floating,)) else v) for k, v in sorted(q.items())},
                      sort_keys=True, default=str)


class Ensemble:
    """Rough realisations ``descs`` (explicit descriptions with rough_std /
    rough_acl / seed) and their incident runs, resident in one batch."""

    def __init__(self, descs, precision=None, directions=('f',), share_incident=True, launch=None,
                 max_slots=None):
        if precision is not None:
            descs = [dict(p, precision=precision) for p in descs]
        self.descs = descs
        self.directions = tuple(directions)
        self.share_incident = share_incident
        inc_desc, self.inc_of = {}, []
        for r, p in enumerate(descs):
            key = incident_key(p) if share_incident else 'r%d' % r
            inc_desc.setdefault(key, configs.incident(p))
            self.inc_of.append(key)
        self.inc_desc = inc_desc
        rough = []
        for p in descs:
            s, up, lo = configs.realise(p)
            rough.append((s, materials.fg_mask(s, up, lo)))
        incs = {}
        for key, q in inc_desc.items():
            s, _, _ = configs.realise(q)
            incs[key] = (s, materials.fg_mask(s, None, None))
        s0 = rough[0][0]
        nslots = len(self.directions) * (len(rough) + len(incs))
        self.batch = Batch(s0.scheme, s0.precision, max_slots=max_slots or nslots)
        if launch is not None:
            self.batch.set_launch(launch[1], launch[2], launch[0])
        self.rough_slot, self.inc_slot = {}, {}
        for d in self.directions:
            for key, (s, fg) in incs.items():
                self.inc_slot[(key, d)] = self.batch.add(s, fg, d)
            for r, (s, fg) in enumerate(rough):
                self.rough_slot[(r, d)] = self.batch.add(s, fg, d)
        self.rough = [s for s, _ in rough]
        self.incident = {k: s for k, (s, _) in incs.items()}
        self.nt = max(s.nt for s in self.rough + list(self.incident.values()))

    @property
    def cell_updates(self):
        """Cell-updates the batch performs (each run over its own nt)."""
        return len(self.directions) * (sum(s.cells * s.nt for s in self.rough) +
                                       sum(s.cells * s.nt for s in self.incident.values()))

    def cell_updates_unshared(self):
        """Cell-updates with one incident run per realisation and direction."""
        inc = sum(self.incident[k].cells * self.incident[k].nt for k in self.inc_of)
        return len(self.directions) * (sum(s.cells * s.nt for s in self.rough) + inc)

    def run(self):
        self.batch.advance(-(-self.nt // 32))
        self.batch.sync()
        return self

    def s_params(self, r):
        """{'S11', 'S21'[, 'S22', 'S12']}: complex [nf] of realisation ``r``."""
        s = self.rough[r]
        N = s.num_cpml
        row0 = window_start(s.scheme, s, 1) - N
        b = self.batch
        out = {}
        for d, (pin, pout, tin, tout) in (('f', (0, 1, 'S11', 'S21')), ('b', (1, 0, 'S22', 'S12'))):
            if d not in self.directions:
                continue
            si = self.inc_slot[(self.inc_of[r], d)]
            sr = self.rough_slot[(r, d)]
            pinc = b.port_power(si, pin, row0, s.pp)
            out[tin] = np.sqrt(b.port_power(sr, pin, row0, s.pp, sub=(si, pin)) / pinc)
            out[tout] = np.sqrt(b.port_power(sr, pout, row0, s.pp) / pinc)
        return out

    def close(self):
        self.batch.close()


def db(z):
    return 20 * np.log10(np.abs(z))


def s_error(a, b):
    """(max |dB difference|, max |phase difference| in degrees) of two
    complex S arrays."""
    ddb = float(np.max(np.abs(db(a) - db(b))))
    dph = float(np.max(np.abs(np.angle(a / b, deg=True))))
    return ddb, dph
