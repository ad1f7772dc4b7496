"""Sidewall roughness profiles and the reference's mask-generation modes.

The reference draws profiles with ``pyspeckle.create_exp_1D`` on numpy's
unseeded global RNG inside a rejection loop (fdtd_funcs.py:147-162), so its
realisations cannot be reproduced.  Here a profile is a stationary Gaussian
AR(1) sequence -- the exact discrete process with autocorrelation
exp(-|k|/Lc) and standard deviation sigma, in cells -- drawn from
``numpy.random.Generator(PCG64(seed))``.  Given the profile arrays, mask and
material maps follow the reference bit for bit (``materials.fg_mask``).

``check_discretization`` restates aux_funcs.py:42-53 (truncate to whole
cells, std, 1/e crossing of the autocorrelation with linear interpolation,
mean).  pyspeckle's ``autocorrelation`` normalisation is not available; the
normalised one-sided autocorrelation of the mean-removed sequence is used.
"""
import os

import numpy as np


def exp_acf_profile(n, sigma_cells, lc_cells, rng):
    """Zero-mean Gaussian sequence of length n with std sigma and ACF exp(-|k|/lc)."""
    e = rng.standard_normal(n)
    x = np.empty(n)
    if n == 0:
        return x
    rho = float(np.exp(-1.0 / lc_cells)) if lc_cells > 0 else 0.0
    inno = sigma_cells * np.sqrt(1.0 - rho * rho)
    x[0] = sigma_cells * e[0]
    for k in range(1, n):
        x[k] = rho * x[k - 1] + inno * e[k]
    return x


def autocorrelation(a):
    a = np.asarray(a, dtype=np.float64)
    a = a - a.mean()
    n = len(a)
    f = np.fft.rfft(a, 2 * n)
    r = np.fft.irfft(f * np.conj(f))[:n]
    return r / r[0] if r[0] != 0 else np.zeros(n)


def check_discretization(test_array, dx, precision=np.float32):
    """(sigma [m], correlation length [m], mean [cells]) of a profile in cells."""
    dist = np.asarray(test_array).astype(int)
    sigma = np.std(dist) * dx
    A = autocorrelation(dist.astype(precision))
    f = A - 1 / np.e
    i = 0
    for i in range(len(f)):
        if f[i] < 0:
            break
    acl = (f[i - 1] / (f[i - 1] - f[i]) + i - 1) * dx if i > 0 else 0.0
    return sigma, acl, float(np.mean(dist))


def make_profile(length_eff, rough_std, rough_acl, dx, stol, atol, mtol, seed=0, max_tries=300000,
                 accept='tolerance'):
    """A profile in cells (rough_std, rough_acl in cells).

    accept='tolerance' keeps the reference's rejection loop and its exception
    after ``max_tries`` draws (fdtd_funcs.py:147-162); accept='first' returns the
    first draw (the statistics are then reported, not enforced).  ``seed`` is an
    int or a ``numpy.random.SeedSequence``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    for i in range(max_tries + 1):
        p = exp_acf_profile(length_eff, rough_std, rough_acl, rng)
        if accept == 'first':
            return p
        sigma, acl, mm = check_discretization(p, dx)
        if ((1 - stol) * rough_std * dx < sigma < (1 + stol) * rough_std * dx and
                (1 - atol) * rough_acl * dx < acl < (1 + atol) * rough_acl * dx and -mtol < mm < mtol):
            return p
    raise Exception('No valid profile for current settings after {:,} profile iterations'.format(max_tries))


def profiles(length, sigma_cells, lc_cells, seed, correlation=3):
    """(upper, lower) profiles for a realisation; correlation as user_config ctype
    (1 equal, 2 opposite, 3 independent -- the s-parameter path always uses 3)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    up = exp_acf_profile(length, sigma_cells, lc_cells, rng)
    if correlation == 1:
        lo = up.copy()
    elif correlation == 2:
        lo = -up
    elif correlation == 3:
        lo = exp_acf_profile(length, sigma_cells, lc_cells, rng)
    else:
        raise Exception('Invalid correlation type selected')
    return up, lo


def profile_names(prof_dir, rough_std, rough_acl, profile_number):
    """fdtd_auto_setup.py:166-170 file stems."""
    tag = '_s{}nm_lc{}nm_r{}'.format(str(round(rough_std * 1e9)), str(round(rough_acl * 1e9)), profile_number)
    return ('./{}/profile_upper{}'.format(prof_dir, tag), './{}/profile_lower{}'.format(prof_dir, tag))


def generate_profiles(mode, length, setup, usr, upper_name='', lower_name='', seed=0):
    """Profiles for mode 'gen' | 'load' | 'smooth' with the reference's file side
    effects (fdtd_funcs.py:165-192): the profiles are saved as .npy in every mode."""
    if mode == 'gen':
        stol, atol = usr['tol_std'] / 100, usr['tol_acl'] / 100
        # Two independent child streams of the realisation's seed: realisation
        # k shares no draw with realisation k + 1 (seeding the lower profile
        # with seed + 1 would repeat the next realisation's upper profile).
        seed_up, seed_lo = np.random.SeedSequence(int(seed)).spawn(2)
        up = make_profile(length, setup.rstd, setup.racl, float(setup.delta_x), stol, atol, 0.01, seed=seed_up,
                          max_tries=usr.get('max_profile_tries', 300000),
                          accept=usr.get('profile_accept', 'tolerance'))
        lo = make_profile(length, setup.rstd, setup.racl, float(setup.delta_x), stol, atol, 0.01, seed=seed_lo,
                          max_tries=usr.get('max_profile_tries', 300000),
                          accept=usr.get('profile_accept', 'tolerance'))
    elif mode == 'load':
        up, lo = np.load(upper_name), np.load(lower_name)
    elif mode == 'smooth':
        up, lo = np.zeros(length), np.zeros(length)
    else:
        raise Exception('Choose a valid roughness profile mode')
    # The reference also saves in 'smooth' mode, where the names are '' and numpy
    # writes a stray '.npy' into the working directory; empty names are skipped.
    for name in (upper_name, lower_name):
        if name and os.path.dirname(name):
            os.makedirs(os.path.dirname(name), exist_ok=True)
    if upper_name:
        np.save(upper_name, up)
    if lower_name:
        np.save(lower_name, lo)
    return up, lo
