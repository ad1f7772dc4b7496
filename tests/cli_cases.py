"""Small reference-style parameter dictionaries for the drop-in driver tests."""
import numpy as np


def small_user(scheme='tm', num_lines=1, rough=True):
    """user_config values shaped like the reference's (user_config.py:29-114)
    but small: a 3 um guide, 5 frequencies, CPML 8."""
    u = dict(eps_rel_bg=1.5 ** 2, eps_rel_fg=3.5 ** 2, sgl_wg_length=3.0e-6, sgl_wg_height=0.2e-6,
             pitch=0.9e-6, num_lines=num_lines, port1_length=1.0e-6, port2_length=1.0e-6,
             rough_std=20.0e-9, rough_acl=200.0e-9, tol_std=10.0, tol_acl=10.0, rough_toggle=rough, ctype=3,
             precision=np.float32, sim_type='s-param', auto_condition=True, feedback_at_n_percent=50,
             profile_number=0, output_dir='Results', prof_dir='rough_profiles', sparam_file='small',
             sparam_fmin=170e12, sparam_fmax=230e12, sparam_num_freqs=5, source_type=2, source_amp=5.0,
             wave_packet_bw=0.5, gauss_pulse_deg=-6, f0=194.8e12, harmonics=1.0, num_flights=2.0,
             points_per_wl=40, num_cpml=8, buffer_xhat=10, ny_clad_wl=1.0, profile_accept='first')
    if scheme == '3d':
        u.update(sgl_wg_width=0.4e-6, sgl_wg_length=2.0e-6, port1_length=0.65e-6, port2_length=0.65e-6,
                 buffer_zhat=8, nx_clad_wl=0.4, ny_clad_wl=0.4, source_condition='TE', pitch=0.3e-6)
        u.pop('buffer_xhat')
    return u
