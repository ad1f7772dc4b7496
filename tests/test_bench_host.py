"""Host-side pieces of bench.py that need no GPU: the CPU baseline's bounded
sample, the committed-traffic lookup and the roofline's kernel naming."""
import bench
from oidt_amd import configs


def test_cpu_baseline_stops_at_the_runs_last_step():
    """A grid small enough to finish its whole run inside the time budget
    stops at the run's own step count (the source waveform has nt entries)."""
    p = dict(configs.config0(), nt=30)
    rate, dt, steps = bench.cpu_baseline(p, 600.0)
    assert steps == 29            # one warm-up step, then the rest of the run
    assert rate > 0 and dt < 600.0


def test_committed_traffic_lookup():
    b, src = bench.committed_traffic('k_3d_tb', 'c3', 4, 2)
    assert src == 'profiles/k_3d_tb_3d_ensemble_16runs_fp32.md'
    assert 2.5e9 < b < 3.5e9
    b, src = bench.committed_traffic('k_3d_e', 'c3', 1, 4)
    assert src == 'profiles/k_3d_e_3d_ensemble_16runs_fp32.md' and b > 1e9
    assert bench.committed_traffic('k_3d_tb', 'c1', 4, 2) == (None, None)


def test_roofline_names_each_schemes_blocked_kernel():
    """Variant 4 reports the blocked kernel of its own scheme, its bytes over
    the launch's steps, and the minimum bytes the launch must move."""
    ms = (1.0, 0.0, 0.1, 0.0)
    for scheme, kname in (('tm', 'k_tmz_tb'), ('te', 'k_tez_tb'), ('3d', 'k_3d_tb')):
        r = bench.roofline_of(scheme, 4, 4, 2, 1000, ms, 'none')
        assert r['kernel'] == kname
        assert r['steps_per_launch'] == 2
        be, bh = bench.algorithmic_bytes(scheme, 4)
        assert r['bytes_per_launch'] == (be + bh) * 1000 * 2
        assert r['min_hbm_bytes_per_launch'] == 2 * (6 if scheme == '3d' else 3) * 4 * 1000
