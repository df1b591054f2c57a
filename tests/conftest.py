import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP kernels)")


def urdf_path(name):
    return os.path.join(GOLDEN, "urdf", name)


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def csd():
    import csdecomp_amd
    return csdecomp_amd


def parse(name):
    import csdecomp_amd as csd
    p = csd.URDFParser()
    assert p.parse_urdf(urdf_path(name))
    return p


@pytest.fixture(scope="session")
def gpu():
    """Initialises the device through the C-ABI; GPU tests fail loudly if it
    cannot (no silent CPU fallback exists)."""
    from csdecomp_amd import capi
    capi.init(0)
    import torch
    dev = torch.device("cuda:0")
    torch.zeros(1, device=dev)  # device memory must be usable
    return dev


BAND = 1e-5  # flags may differ only within 1e-5 m of contact (FP64 distance)


def assert_flags_match(flags, ref_flags, sd, what=""):
    flags = np.asarray(flags).astype(bool)
    ref_flags = np.asarray(ref_flags).astype(bool)
    outside = np.abs(sd) >= BAND
    bad = np.nonzero((flags != ref_flags) & outside)[0]
    assert bad.size == 0, f"{what}: {bad.size} flag mismatches outside the contact band, e.g. {bad[:10]}"
    # the FP64 sign decides every configuration outside the band
    assert np.all(flags[outside] == (sd[outside] > 0)), f"{what}: flags disagree with FP64 sign"
