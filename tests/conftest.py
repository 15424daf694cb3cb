import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def core_golden():
    return load_golden("delhommeau_core.npz")


@pytest.fixture(scope="session")
def bm_golden():
    return load_golden("build_matrices.npz")


@pytest.fixture(scope="session")
def pp_golden():
    return load_golden("post_processing.npz")


@pytest.fixture(scope="session")
def nemoh_golden():
    return load_golden("nemoh2_known_answers.npz")


@pytest.fixture(scope="session")
def orc():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def default_table(orc, core_golden):
    """The default (scaled_nemoh3, 676 x 372, 1001 points) table from the
    oracle, checked bit for bit against the reference table checksum."""
    import hashlib
    print("[fixture] oracle default table (676 x 372 x 1001 points)", flush=True)
    r = orc.r_spacing(676, 100.0, 1)
    z = orc.z_spacing(372, -251.0, 1)
    tab = orc.tabulation(r, z, 1001)
    sha = hashlib.sha256(np.asfortranarray(tab).tobytes(order="F")).hexdigest()
    assert sha == str(core_golden["table_scaled_sha256"])
    return r, z, tab


@pytest.fixture(scope="session")
def legacy_table(orc, core_golden):
    import hashlib
    print("[fixture] oracle legacy table", flush=True)
    r = orc.r_spacing(328, 100.0, 0)
    z = orc.z_spacing(46, -16.0, 0)
    tab = orc.tabulation(r, z, 251)
    sha = hashlib.sha256(np.asfortranarray(tab).tobytes(order="F")).hexdigest()
    assert sha == str(core_golden["table_legacy_sha256"])
    return r, z, tab


def rel_err(a, b):
    """Normwise relative error ||a - b|| / ||b||."""
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def elem_ok(a, b, rtol, floor):
    """Elementwise |a - b| <= rtol * max(|b|, floor * max|b|)."""
    a, b = np.asarray(a), np.asarray(b)
    scale = np.maximum(np.abs(b), floor * np.max(np.abs(b)))
    return np.all(np.abs(a - b) <= rtol * scale)
