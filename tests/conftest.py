"""Shared fixtures: small generated grids and their atlases, built once per session."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "jax-rans_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

from jaxrans_amd.geometry.airfoil import karman_trefftz, naca4  # noqa: E402
from jaxrans_amd.mesh.assemble import assemble  # noqa: E402
from jaxrans_amd.mesh.atlas import assign_conditions, build_atlas  # noqa: E402
from jaxrans_amd.mesh.generate import c_grid, o_grid  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: longer CPU runs")


def pytest_collection_modifyitems(config, items):
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(pytest.mark.timeout(300))


_CACHE: dict = {}


def _cached(key, make):
    if key not in _CACHE:
        _CACHE[key] = make()
    return _CACHE[key]


def o_atlas(n_around=48, n_normal=24, far=15.0, wall="slip_wall", spacing=2e-3):
    def make():
        X, Y = o_grid(naca4("0012", sharp=False), n_around, n_normal, spacing, far)
        mesh = assign_conditions(assemble([(X, Y)]), {"wall": wall, "farfield": "far_field"})
        return build_atlas(mesh)
    return _cached(("o", n_around, n_normal, far, wall, spacing), make)


def c_atlas(n_air=48, n_wake=12, n_normal=24, far=15.0, spacing=1e-3):
    def make():
        X, Y = c_grid(karman_trefftz(10.0, (-0.08, 0.04), 129), n_air, n_wake, n_normal, spacing, far)
        mesh = assign_conditions(assemble([(X, Y)]), {"wall": "no_slip_wall", "farfield": "far_field"})
        return build_atlas(mesh)
    return _cached(("c", n_air, n_wake, n_normal, far, spacing), make)


def perturbed(U0: np.ndarray, amplitude: float = 1e-3) -> np.ndarray:
    """The freestream times ``1 + a sin(k)`` over the flat index: a non-trivial
    state (the micro-benchmark input of the reference's scripts/benchmark.py:93-101)."""
    return U0 * (1.0 + amplitude * np.sin(np.arange(U0.size)).reshape(U0.shape))


@pytest.fixture(scope="session")
def euler_atlas():
    return o_atlas()


@pytest.fixture(scope="session")
def sa_atlas():
    return c_atlas()
