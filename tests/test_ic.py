"""The C++ initial conditions and the batch sampler (libccamd, no device needed) against the
independent numpy restatements in tests/ic_restated.py, plus the properties the benchmark ICs are
defined by (solenoidal velocity, rms Mach, spectrum, plasma beta, the sampler's ranges).

The restatements use numpy's transcendental functions where the C++ uses libm, so the MHD ICs
agree to a few ulp (rtol 1e-13); the draws from std::mt19937 and the sampler's arithmetic are exact.
"""
import numpy as np
import pytest

from core_collapse_amd import ic
from core_collapse_amd.mhd2d import Params as M2Params
from core_collapse_amd.mhd3d import Params as M3Params
from tests import ic_restated as R

RTOL = 1e-13


def _close(a, b, name):
    a, b = np.asarray(a).reshape(-1), np.asarray(b).reshape(-1)
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= RTOL * scale, (name, np.abs(a - b).max(), scale)


def test_mhd2d_reference_ic_matches_restatement():
    p = M2Params(Nx=40, Ny=56)
    got = ic.mhd2d_reference_ic(p)
    want = R.mhd2d_reference_ic(p)
    for f in want:
        _close(got[f], want[f], f)


@pytest.mark.parametrize("knobs", [{}, {"seed": 7, "contrast": 37.5, "width": 0.13, "P0": 0.2, "vamp": 3e-3,
                                        "B0": 0.08}])
def test_collapse2d_ic_matches_restatement(knobs):
    p = ic.collapse2d_params(48, Ny=36)
    got = ic.collapse2d_ic(p, **knobs)
    want = R.collapse2d_ic(p, **knobs)
    for f in want:
        _close(got[f], want[f], f)
    # the benchmark's defining values: peak rho -> contrast at the centre, uniform Bx, p = P0
    c = knobs.get("contrast", 10.0)
    assert abs(got["rho"].max() - c) < 0.1 * c
    assert np.all(got["Bx_f"] == knobs.get("B0", 0.05)) and np.all(got["By_f"] == 0.0)


def test_mhd3d_reference_ic_matches_restatement():
    p = M3Params(Nx=12, Ny=10, Nz=14)
    got = ic.mhd3d_reference_state(p)
    want = R.mhd3d_reference_state(p)
    for f in want:
        _close(got[f], want[f], f)


def test_turbulence3d_ic_matches_restatement_and_is_solenoidal():
    p = ic.turbulence3d_params(32)
    got = ic.turbulence3d_ic(p, seed=11, mach=1.0, beta=1.0)
    want = R.turbulence3d_ic(p, seed=11)
    for f in want:
        _close(got[f], want[f], f)
    # solenoidal: the spectral divergence of v vanishes to rounding
    rho = got["rho"]
    u, v, w = got["mx"] / rho, got["my"] / rho, got["mz"] / rho
    k = np.fft.fftfreq(32, 1.0 / 32)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    div = KX * np.fft.fftn(u) + KY * np.fft.fftn(v) + KZ * np.fft.fftn(w)
    assert np.abs(div).max() <= 1e-12 * np.abs(np.fft.fftn(u)).max() * 32
    # rms Mach 1 at sound speed 1 (rho = 1, p = 1/gamma), beta = 2 p / B0^2 = 1
    assert np.allclose(rho, 1.0)
    assert abs(np.sqrt(np.mean(u * u + v * v + w * w)) - 1.0) < 1e-12
    B0 = got["Bz_f"][0, 0, 0]
    assert np.all(got["Bz_f"] == B0) and abs(2.0 * (1.0 / p.gamma) / (B0 * B0) - 1.0) < 1e-14
    # all power in 1 <= |k| <= 8
    kk = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    pw = np.abs(np.fft.fftn(u)) ** 2 + np.abs(np.fft.fftn(v)) ** 2 + np.abs(np.fft.fftn(w)) ** 2
    assert pw[(kk < 1) | (kk > 8)].sum() <= 1e-20 * pw.sum()


def _sample_restated(seed, n, nmin, nmax):
    """The sampler from its definition: per instance four std::mt19937(seed) draws u = (r + 0.5)/2^32
    in the order (grid, contrast, eta, B0); N = round(nmin (nmax/nmin)^u) half away from zero."""
    u = (R.mt19937_raw(seed, 4 * n).astype(np.float64) + 0.5) / 4294967296.0
    uN, uc, ue, uB = u[0::4], u[1::4], u[2::4], u[3::4]
    N = np.floor(nmin * np.power(nmax / nmin, uN) + 0.5).astype(int)
    return N, 5.0 + 45.0 * uc, 1e-5 + (1e-3 - 1e-5) * ue, 0.01 + 0.09 * uB


def test_batch_sampler_matches_its_definition():
    s = ic.collapse2d_batch_sample(64, seed=42)
    N, c, eta, B0 = _sample_restated(42, 64, 128, 1024)
    assert [e["N"] for e in s] == list(N)
    assert [e["ic"]["contrast"] for e in s] == list(c)
    assert [e["eta"] for e in s] == list(eta)
    assert [e["ic"]["B0"] for e in s] == list(B0)
    assert [e["ic"]["seed"] for e in s] == [42 + i for i in range(64)]
    d = ic.collapse2d_batch_sample(1)[0]["ic"]
    assert d["width"] == 0.1 and d["P0"] == 0.1 and d["vamp"] == 1e-3


def test_batch_sampler_ranges_and_log_uniform_grids():
    s = ic.collapse2d_batch_sample(4000, seed=3)
    N = np.array([e["N"] for e in s])
    assert N.min() >= 128 and N.max() <= 1024
    # log-uniform: each octave [128,256), [256,512), [512,1024] holds about a third
    frac = np.histogram(np.log2(N), bins=[7, 8, 9, 10.01])[0] / len(N)
    assert np.all(np.abs(frac - 1 / 3) < 0.03), frac
    for key, lo, hi in (("eta", 1e-5, 1e-3),):
        v = np.array([e[key] for e in s])
        assert v.min() >= lo and v.max() <= hi and abs(v.mean() - 0.5 * (lo + hi)) < 0.03 * (hi - lo)
    for key, lo, hi in (("contrast", 5.0, 50.0), ("B0", 0.01, 0.1)):
        v = np.array([e["ic"][key] for e in s])
        assert v.min() >= lo and v.max() <= hi and abs(v.mean() - 0.5 * (lo + hi)) < 0.03 * (hi - lo)
    assert ic.collapse2d_batch_sample(8, seed=3) == s[:8]   # deterministic prefix
