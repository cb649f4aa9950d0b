"""The oracle's building blocks against known answers and closed forms, and
the product library's host-side tables against the oracle's (both run on
the CPU: these functions of the C ABI need no GPU)."""
import math

import numpy as np
import pytest

from molflow_sources_amd import model as M
from oracle import mfcpu


# Random123 known-answer vectors of Philox4x32-10 (kat_vectors, Salmon et al. SC'11)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(oracle_lib, ctr, key, expect):
    assert tuple(mfcpu.philox(ctr, key)) == expect


def test_maxwell_cdf_closed_form(oracle_lib):
    # CDFGeneration.cpp:55-83 at 1 K: P(v) = 1 - exp(-x)(1+x), x = v^2/(2a^2),
    # a^2 = kT/m, sampled on 100 points over [0, 4 v_p).
    for mass in (2.0, 28.0, 131.0):
        v, p = mfcpu.maxwell_cdf(mass)
        a2 = 1.38e-23 * 1.0 / (mass * 1.67e-27)
        vp = math.sqrt(2 * 8.3144621 / (mass / 1000.0))
        assert v[0] == 0 and p[0] == 0
        np.testing.assert_allclose(np.diff(v), 4 * vp / 100, rtol=1e-9)
        x = v ** 2 / (2 * a2)
        np.testing.assert_allclose(p, 1 - np.exp(-x) * (1 + x), rtol=1e-12, atol=1e-15)
        assert np.all(np.diff(p) > 0)
        # flux-weighted (wall collision) speed distribution ~ v^3 exp(-v^2/2a^2):
        # its mean is (3/4) sqrt(2 pi a^2); check the table's inverse CDF
        u = (np.arange(200000) + 0.5) / 200000
        sp = np.interp(u, p, v)
        ok = u < p[-1]
        mean = np.mean(sp[ok]) * ok.mean() + (1 - ok.mean()) * v[-1]
        assert abs(mean / (0.75 * math.sqrt(2 * math.pi * a2)) - 1) < 0.01


def test_gpu_library_tables_match_oracle(oracle_lib):
    from molflow_sources_amd import simulation as S
    for mass in (4.0, 28.0):
        v1, p1 = mfcpu.maxwell_cdf(mass)
        v2, p2 = S.maxwell_cdf(mass)
        assert np.array_equal(v1, v2) and np.array_equal(p1, p2)
    for content, window in (("1e-4,1e-4,1e-3", 1e-4), ("2e-3,1e-3,3e-2", 1e-3), ("0.25e-3,0.5e-3,49.75e-3", 0.5e-3),
                            ("0,.1,3", 0.1), ("0.5", 1.0), ("1,1,0.5", 0.1)):
        assert np.array_equal(mfcpu.parse_user_moment(content, window), S.parse_user_moment(content, window))
    x, y = [0.0, 1e-3, 1.000001e-3], [1.0, 1.0, 0.0]
    a = mfcpu.integrated_desorption(x, y, False, False, 0.05)
    b = S.integrated_desorption(x, y, False, False, 0.05)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("content,window,count", [
    # moment counts stored in the reference fixtures (results nb = 1 + moments)
    ("1e-4,1e-4,1e-3", 1e-4, 9), ("2e-3,1e-3,3e-2", 1e-3, 28),   # fixture 02: 38 slots
    ("1e-3,1e-3,4e-2", 1e-3, 39),                                  # fixture 07: 40 slots
    ("1e-3,1e-3,1e-1", 1e-3, 99),                                  # fixture 08: 100 slots
    ("0,.1,3", 0.1, 30),                                           # fixture 02c: 31 slots
    ("1e-5,2e-5,9e-5", 2e-5, 5),                                   # fixture 09: 6 slots
    ("0.25e-3,0.5e-3,49.75e-3", 0.5e-3, 99),                       # accumulation drops the end point
    ("0.25e-3,0.5e-3,49.8e-3", 0.5e-3, 100),
    ("1e-3,1e-4,2e-3", 1e-3, 0),                                   # overlapping windows rejected
    ("0.3", 1e-3, 1),
])
def test_parse_user_moment_known_answers(oracle_lib, content, window, count):
    assert len(mfcpu.parse_user_moment(content, window)) == count
    assert len(M.parse_user_moment(content, window)) == count


def test_integrated_desorption_rectangular_pulse(oracle_lib):
    # Q = 2 mbar.l/s for 1 ms then 0: cumulative = Q * t * 0.1 Pa.m3 (IDGeneration.cpp:107-110)
    t, c = mfcpu.integrated_desorption([0.0, 1e-3, 1.000001e-3], [2.0, 2.0, 0.0], False, False, 0.05)
    assert t[0] == 0 and c[0] == 0
    k = np.searchsorted(t, 1e-3)
    assert abs(c[k] - 2.0 * 1e-3 * 0.1) < 1e-15
    assert abs(c[-1] - (2.0 * 1e-3 * 0.1 + 0.5 * 2.0 * 1e-9 * 0.1)) < 1e-15
    assert abs(t[-1] - 0.05) < 1e-15


def test_integrated_desorption_catalog_dirac(oracle_lib):
    # the reference's parameter catalog dirac_burst.csv: (1e-6, 1e6), (1.000001e-6, 0)
    t, c = mfcpu.integrated_desorption([1e-6, 1.000001e-6], [1e6, 0.0], False, False, 1.0)
    # constant 1e6 from 0 to 1e-6, then a 10-step trapezoid down to 0 over 1e-12 s
    expect = (1e6 * 1e-6 + 0.5 * 1e6 * 1e-12) * 0.1
    assert abs(c[-1] - expect) < 1e-9 * expect


def test_integrated_desorption_log_modes_integrate_exactly(oracle_lib):
    # lin-lin integral of a linear ramp is exact; log-log integral of a power law is exact
    t, c = mfcpu.integrated_desorption([0.0, 1.0], [0.0, 2.0], False, False, 1.0)
    assert abs(c[-1] - 0.1 * 1.0) < 1e-12
    x, y = [1.0, 10.0], [1.0, 100.0]            # y = x^2 on a log-log plot
    t, c = mfcpu.integrated_desorption(x, y, True, True, 10.0)
    exact = 0.1 * (1.0 * 1.0 + (1000.0 - 1.0) / 3.0)   # const 1 on [0,1] + int_1^10 x^2
    assert abs(c[-1] / exact - 1) < 1e-9
