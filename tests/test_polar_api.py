"""The batched polar's interface (``src/jaxrans/polar.py:43-121``): the call
order of ``run_polar`` and the ``Polar`` record's table and files, on the CPU
with stand-in results (the device batch itself is covered by the GPU tests)."""

from __future__ import annotations

import inspect
import json
from types import SimpleNamespace

import numpy as np

from jaxrans_amd.simulation import SETTLED_CL_SPREAD, Polar, Timing, run_polar


def test_run_polar_takes_the_reference_positional_order():
    params = list(inspect.signature(run_polar).parameters.values())
    assert [p.name for p in params[:5]] == ["case", "alphas", "grid", "devices", "progress"]
    extras = {p.name: p.kind for p in params[5:]}
    assert extras == {"machs": inspect.Parameter.KEYWORD_ONLY, "reynolds": inspect.Parameter.KEYWORD_ONLY}


class _Stub:
    def __init__(self, cl, cd, cm, cdf, hist, cl_hist):
        self.forces = SimpleNamespace(cl=cl, cd=cd, cm=cm, cd_friction=cdf)
        self.history = SimpleNamespace(orders_dropped=float(np.log10(hist[0] / hist[-1])), n_iters=len(hist))
        self.cl_spread = float(np.ptp(np.asarray(cl_hist)[3 * len(cl_hist) // 4:]))
        self.written = None

    def write(self, directory):
        directory.mkdir(parents=True, exist_ok=True)
        self.written = directory
        return directory


def test_polar_table_and_files(tmp_path):
    a = _Stub(0.1, 0.01, -0.02, 0.004, [1.0, 1e-3], [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.1])
    b = _Stub(0.5, 0.02, -0.05, 0.005, [2.0, 2e-6], [0.3, 0.4, 0.5, 0.45, 0.5, 0.52, 0.47, 0.5])
    timing = Timing(total=2.0, first_block=1.0, iterations=40, block_iters=20, cells=100, cases=2)
    polar = Polar(case=SimpleNamespace(name="c", raw={"k": 1}), alphas=np.array([-1.0, 2.5]),
                  results=[a, b], timing=timing)
    t = polar.table()
    assert t.shape == (2, 9)
    assert np.allclose(t[0], [-1.0, 0.1, 0.01, 0.004, -0.02, 3.0, 2, 0.0, 1.0])
    assert t[1, 5] == np.log10(1e6) and t[1, 7] == np.ptp([0.5, 0.47]) and t[1, 8] == float(0.03 <= SETTLED_CL_SPREAD)
    out = polar.write(tmp_path / "p")
    back = np.loadtxt(out / "polar.dat")
    assert np.allclose(back[:, [0, 1, 2, 6]], t[:, [0, 1, 2, 6]])
    meta = json.loads((out / "polar_metadata.json").read_text())
    assert meta["alphas"] == [-1.0, 2.5] and meta["case"] == "c"
    assert a.written.name == "alpha_-1.00" and b.written.name == "alpha_+2.50"
