import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import __graft_entry__ as _entry  # noqa: E402
import oracle as _oracle  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def ND():
    return _entry.load()


@pytest.fixture(scope="session")
def oracle():
    return _oracle


def oracle_for(nw):
    """The oracle network for a product Network (same graph, kinds and kind assignment,
    and the same ExtMap where the network has external inputs)."""
    on = _oracle.OracleNet(nw.nv(), nw.g.src, nw.g.dst, nw.kinds, nw.vertex_kind, nw.edge_kind)
    if nw.info.lastidx_extbuf > 0:
        on.set_extmap(nw.extmap())
    return on


def fd_jvp(on, u, v, p, h=1e-3):
    """Richardson-extrapolated central difference of the oracle RHS along v: the
    directional derivative J(u) v with an O(h^4) truncation error."""
    def d(hh):
        return (on.rhs(u + hh * v, p) - on.rhs(u - hh * v, p)) / (2 * hh)
    return (4 * d(h / 2) - d(h)) / 3


def scaled_error(du, du_ref, S):
    """|du - du_ref| / max(|du_ref|, S): S is the sum of the magnitudes of each element's terms."""
    den = np.maximum(np.abs(du_ref), S)
    den = np.where(den == 0, 1.0, den)
    return np.abs(du - du_ref) / den


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")
