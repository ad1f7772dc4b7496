"""Serving on the device: continuous batching admits requests while others run;
every result is bit-identical to the same request solved alone."""
import numpy as np
import pytest

from oidt_amd import configs, serve
from oidt_amd.solver import Solver


def _requests():
    descs = []
    for i, (scheme, extra) in enumerate([('tm', 0), ('tm', 30), ('3d', 0), ('tm', 60), ('3d', 4), ('tm', 90)]):
        if scheme == 'tm':
            p = configs.small('tm', 'float32', nt=300 + 64 * i)
            p['nx'] += extra
            p['p2'] += extra
        else:
            p = configs.small('3d', 'float32', nt=200 + 32 * i)
            p['nz'] += extra
            p['p2'] += extra
        p.update(rough_std=(5 + 2 * i) * 1e-9, rough_acl=40e-9, seed=100 + i)
        descs.append(p)
    return descs


FUSED = {'tm': (3, 16, 0), '3d': (3, 4, 0)}
BLOCKED = {'tm': (4, 8, 0), '3d': (1, 4, 0)}
BLOCKED_3D = {'tm': (4, 4, 0), '3d': (4, 2, 0)}


@pytest.mark.gpu
@pytest.mark.parametrize('budget_runs,launch', [(100, None), (2, None), (100, FUSED), (2, BLOCKED),
                                                (100, BLOCKED_3D)])
def test_results_identical_alone_and_in_mix(budget_runs, launch):
    """Also with the fused step forms, whose runs alternate between two field
    sets and share per-wave tile lists, and the blocked steps (TMz 8 or 4,
    3D 2 steps per launch): a run admitted while others are half way through
    still gives the phasors it gives alone."""
    descs = _requests()
    reqs = serve.make_requests(descs, [0.0, 0.002, 0.004, 0.006, 0.02, 0.03])
    be = serve.DeviceBackend(max_slots=8, launch=launch)
    fp = max(be.footprint(configs.realise(d)[0]) for d in descs)
    srv = serve.Server(be, hbm_budget=budget_runs * fp, chunk_group=2)
    done = srv.run(reqs)
    be.close()
    assert all(r.error is None for r in done)
    assert srv.peak_bytes <= budget_runs * fp
    if budget_runs == 2:
        assert srv.peak_resident < len(reqs)      # admission waited for memory
    else:
        assert srv.peak_resident >= 3            # requests overlapped on the device
    for r in done:
        print('request', r.rid, r.setup.scheme, r.setup.shape, r.setup.nt, 'latency %.3f' % r.latency, flush=True)
        alone = Solver(r.setup, r.fg)
        if launch:
            variant, xchunk, grid = launch[r.setup.scheme]
            alone.set_launch(xchunk, grid, variant)
        alone.run()
        ref = alone.dft()
        alone.close()
        for k, z in ref.items():
            assert np.array_equal(z, r.result[k]), (r.rid, k)
