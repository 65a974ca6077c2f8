"""Orbits read from sampled positions (an orbit file) and interpolated
linearly, as lisatools' orbit classes are with ``linear_interp_setup``
(src/gbgpu/gbgpu.py:86-101, :390-397).  Pinned against numpy's ``interp``
(the oracle's TableOrbits), the exact samples at the nodes and the
second-order chord bound of linear interpolation."""

import numpy as np
import pytest

from gbgpu_amd import GBGPU, EqualArmlengthOrbits, InterpolatedOrbits
from gbgpu_amd.orbits import orbit_table
from gbgpu_amd.utils.constants import AU_SI, LISA_FM
from oracle import fastgb as F

T4 = 4 * 31_557_600.0


def sampled(dt=1e4, span=T4, orbits=None):
    a = orbits or EqualArmlengthOrbits()
    t = np.arange(int(np.floor(span / dt)) + 1) * dt
    x = np.stack([a.get_pos(t, s) for s in (1, 2, 3)], axis=1)
    return t, x


def test_nodes_are_exact_and_match_numpy_interp():
    t, x = sampled()
    o = InterpolatedOrbits(t, x)
    assert o.dt_base == 1e4
    for s in (1, 2, 3):
        assert np.array_equal(o.get_pos(t, s), x[:, s - 1])
    r = F.TableOrbits(t, x)
    tt = np.random.default_rng(1).uniform(0, t[-1], 50_000)
    for s in (1, 2, 3):
        assert np.abs(o.get_pos(tt, s) - r.get_pos(tt, s)).max() < 1e-3  # metres of 1.5e11 (a few ulp)
    # non-uniform samples (binary-search path) agree with numpy too
    keep = np.sort(np.unique(np.r_[0, np.random.default_rng(2).choice(len(t), 3000, replace=False), len(t) - 1]))
    o2, r2 = InterpolatedOrbits(t[keep], x[keep]), F.TableOrbits(t[keep], x[keep])
    assert o2.dt_base is None
    for s in (1, 2, 3):
        assert np.abs(o2.get_pos(tt, s) - r2.get_pos(tt, s)).max() < 1e-3


def test_interpolation_error_is_the_chord_bound():
    """|P_lin - P| <= dt^2/8 max|P''| with |P''| <= AU (2 pi fm)^2 (1 + 4 ec)
    for the analytic orbits, and the midpoint error reaches at least half of
    it (the samples are interpolated, not re-evaluated)."""
    dt = 2e4
    t, x = sampled(dt=dt, span=0.5 * T4)
    o = InterpolatedOrbits(t, x)
    a = EqualArmlengthOrbits()
    bound = dt**2 / 8 * AU_SI * (2 * np.pi * LISA_FM) ** 2 * (1 + 4 * a.ec)
    mid = t[:-1] + dt / 2
    for s in (1, 2, 3):
        err = np.linalg.norm(o.get_pos(mid, s) - a.get_pos(mid, s), axis=-1)
        assert err.max() <= bound
        assert err.max() >= 0.5 * bound
        # a chord of a convex orbit lies inside it: the interpolated point is
        # closer to the Sun than the true one
        assert np.all(np.linalg.norm(o.get_pos(mid, s), axis=-1) < np.linalg.norm(a.get_pos(mid, s), axis=-1))


def test_configure_resamples_onto_a_uniform_grid(tmp_path):
    t, x = sampled(dt=1e4, span=0.1 * T4)
    o = InterpolatedOrbits(t, x)
    o.configure(linear_interp_setup=True, dt=3e4)
    assert o.dt_base == 3e4 and o.linear_interp_setup
    assert np.array_equal(o.t_base, t[::3][: len(o.t_base)])
    for s in (1, 2, 3):
        assert np.array_equal(o.get_pos(o.t_base, s), x[::3, s - 1][: len(o.t_base)])
    # what GBGPU's orbit setter calls (no dt) keeps the samples
    o2 = InterpolatedOrbits(t, x)
    GBGPU.__new__(GBGPU).orbits = o2
    assert o2.linear_interp_setup and np.array_equal(o2.t_base, t)


def test_file_round_trip_and_errors(tmp_path):
    t, x = sampled(dt=1e5, span=0.1 * T4)
    o = InterpolatedOrbits(t, x, armlength=2.4e9)
    path = tmp_path / "orbits.npz"
    o.save(path)
    o2 = InterpolatedOrbits.from_file(path)
    assert o2.armlength == 2.4e9 and np.array_equal(o2.t_base, t)
    assert InterpolatedOrbits.from_file(path, armlength=2.5e9).armlength == 2.5e9
    with pytest.raises(ValueError, match="outside"):
        o.get_pos(np.array([t[-1] + 1.0]), 1)
    with pytest.raises(ValueError, match="outside"):
        o.get_pos(np.array([-1.0]), 2)
    with pytest.raises(ValueError, match="1, 2 or 3"):
        o.get_pos(t[:3], 4)
    with pytest.raises(ValueError, match="shape"):
        InterpolatedOrbits(t, x[:, :2])
    with pytest.raises(ValueError, match="increasing"):
        InterpolatedOrbits(t[::-1], x)


def test_observation_time_checked_against_the_samples():
    """run_wave refuses T beyond the last sample, as the reference does
    against its orbit file (src/gbgpu/gbgpu.py:181-184); the orbit table
    built from the samples matches the one built from the exact orbits
    to the interpolation error."""
    t, x = sampled(dt=1e4, span=1.2 * F.YRSID_SI)
    gb = GBGPU.__new__(GBGPU)
    gb.orbits = InterpolatedOrbits(t, x)
    with pytest.raises(ValueError, match="orbital information"):
        gb._prepare(*np.ones((9, 1)), (), 128, 1.3 * F.YRSID_SI, 15.0, 1)
    tab = orbit_table(gb.orbits, F.YRSID_SI, 256)
    ref = orbit_table(EqualArmlengthOrbits(), F.YRSID_SI, 256)
    assert np.abs(tab[1:4] - ref[1:4]).max() < 1e5  # guiding centre, metres
    assert np.abs(tab[13:] - ref[13:]).max() < 1e-6  # r_ij, units of armlength
