import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(TESTS, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running")


def example(name):
    """Reference example graphs (examples/*.graph), stored gzipped under tests/golden/."""
    from kahip_amd import graphs
    return graphs.read_metis(os.path.join(GOLDEN, name + ".graph.gz"))


@pytest.fixture(scope="session")
def rgg15():
    return example("rgg_n_2_15_s0")


@pytest.fixture(scope="session")
def delaunay15():
    return example("delaunay_n15")


@pytest.fixture(scope="session")
def weighted132():
    return example("example_weighted")


def manual_graph():
    """The 5-node graph of the reference's library demo (misc/example_library_call/
    interface_test.cpp:18-24)."""
    xadj = np.array([0, 2, 5, 7, 9, 12], np.int32)
    adjncy = np.array([1, 4, 0, 2, 4, 1, 3, 2, 4, 0, 1, 3], np.int32)
    return xadj, adjncy, None, None


def random_graph(n, p, seed, weighted=False):
    """Small Erdos-Renyi test graph (symmetric, no self-loops)."""
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(n, 1)
    keep = rng.random(len(iu[0])) < p
    u, v = iu[0][keep], iu[1][keep]
    w = rng.integers(1, 9, len(u)) if weighted else np.ones(len(u), np.int64)
    uu = np.concatenate([u, v])
    vv = np.concatenate([v, u])
    ww = np.concatenate([w, w])
    order = np.lexsort((vv, uu))
    uu, vv, ww = uu[order], vv[order], ww[order]
    xadj = np.zeros(n + 1, np.int64)
    np.add.at(xadj, uu + 1, 1)
    xadj = np.cumsum(xadj).astype(np.int32)
    vw = rng.integers(1, 5, n).astype(np.int32) if weighted else None
    return xadj, vv.astype(np.int32), vw, (ww.astype(np.int32) if weighted else None)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
