import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rte-rrtmgp-cpp_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

AVAILABLE = ["h2o", "co2", "o3", "n2o", "ch4"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    import build_oracle  # noqa: E402  (oracle recipe, test infrastructure)
    src = os.path.join(ROOT, "oracle", "rrtmgp_oracle.cpp")
    if (not os.path.exists(build_oracle.LIB)
            or os.path.getmtime(build_oracle.LIB) < os.path.getmtime(src)):
        build_oracle.build()


@pytest.fixture(scope="session")
def syn():
    from rrtmgp_amd import synthetic
    return synthetic


@pytest.fixture(scope="session")
def kd_lw(syn):
    return syn.make_kdist(True, syn.SEED_KDIST_LW)


@pytest.fixture(scope="session")
def kd_sw(syn):
    return syn.make_kdist(False, syn.SEED_KDIST_SW)


@pytest.fixture(scope="session")
def lut_lw(syn):
    return syn.make_cloud_lut(True, syn.SEED_CLOUD_LW)


@pytest.fixture(scope="session")
def lut_sw(syn):
    return syn.make_cloud_lut(False, syn.SEED_CLOUD_SW)


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def okd_lw(O, kd_lw):
    return O.Kdist(kd_lw, AVAILABLE)


@pytest.fixture(scope="session")
def okd_sw(O, kd_sw):
    return O.Kdist(kd_sw, AVAILABLE)


@pytest.fixture(scope="session")
def atm_c0(syn):
    """Config-0 shaped clear-sky profiles, reduced to 16 columns for test speed."""
    return syn.make_atmosphere(syn.SEED_ATMOS[0], 16, 60)


@pytest.fixture(scope="session")
def atm_cloudy(syn):
    return syn.make_atmosphere(syn.SEED_ATMOS[1], 24, 128, cloudy=True)


class Err(float):
    """A float error measure whose repr names the worst element (shown by failing asserts)."""

    def __new__(cls, value, info=""):
        obj = float.__new__(cls, value)
        obj.info = info
        return obj

    def __repr__(self):
        return f"{float(self):.3e} [{self.info}]"


def rel_err(a, b, floor=0.0):
    """max |a - b| / max(|b|, floor), carrying the index and both values of the worst element."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if not a.size:
        return Err(0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.abs(a - b) / np.maximum(np.abs(b), floor)
    e = np.where(a == b, 0.0, np.where(np.isnan(e), np.inf, e))
    i = np.unravel_index(int(np.argmax(e)), e.shape)
    return Err(e[i], f"at {tuple(int(x) for x in i)} of {e.shape}: got {a[i]!r} want {b[i]!r}")
