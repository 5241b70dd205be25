"""Shared fixtures.  `gpu` tests need an MI355X (gfx950) and call the HIP
library through its C-ABI; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU")


def _ensure_built():
    from wildfires_amd import _abi, build
    if not os.path.exists(_abi.LIB_PATH):
        build.build()
    from oracle import build as obuild
    if not os.path.exists(obuild.LIB):
        obuild.build()


_ensure_built()


@pytest.fixture
def workdir(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    return tmp_path


def short_dir():
    """A fresh directory with a short absolute path: the run's output directory
    must fit the 64-byte save_path of the Parameters structure."""
    import tempfile
    return tempfile.mkdtemp(prefix="wf", dir="/tmp")


def make_params(tmp_path, name, **over):
    from wildfires_amd import configs, api
    path = os.path.join(str(tmp_path), name + ".txt")
    configs.write_parameter_file(path, name, **over)
    return api.read_parameters(path, output_root=short_dir())


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    if den == 0.0:
        return float(np.linalg.norm(a - b))
    return float(np.linalg.norm(a - b) / den)


def field_errors(params, a, b):
    """Relative L2 error per field (u, v, w, T, Y) of state vectors a vs b."""
    return {n: rel_l2(x, y) for n, x, y in zip("uvwTY", params.split(a), params.split(b))}


def perturbed_state(params, seed=7, amp=0.3):
    """Initial state plus smooth random perturbations of u, v, w, T so every
    term of the right-hand side is non-zero."""
    from wildfires_amd import api
    y0, p0 = api.initial_conditions(params)
    rng = np.random.default_rng(seed)
    u, v, w, T, Y = params.split(y0)
    shape = u.shape
    for f, scale in ((u, amp), (v, amp), (w, amp), (T, 40.0 * amp)):
        f += scale * rng.standard_normal(shape)
    cut = params.cut_nodes.reshape(shape[0], shape[1])
    kz = np.arange(shape[2])[None, None, :]
    dead = kz < cut[:, :, None]
    u[dead] = params.p.u_dead_nodes
    v[dead] = params.p.v_dead_nodes
    w[dead] = params.p.w_dead_nodes
    T[dead] = params.p.T_dead_nodes
    p0 += rng.standard_normal(p0.shape)
    return y0, p0
