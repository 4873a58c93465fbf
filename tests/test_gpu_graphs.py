"""The root searches replay captured graphs of pass + step kernels (one per grid, precision and
search kind, captured once a chunk has been launched TPSC_GRAPH_AFTER times; 0 here, so the
first use captures).  Several solves of different points in one context (graph replays with
eager stages and host waits between them) must give, bit for bit, what the same kernels give
when launched one by one (TPSC_NO_GRAPHS=1, run in a child process)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POINTS = [(0.9, 0.15, 3.0, -0.1), (0.85, 0.2, 2.5, -0.2), (0.95, 0.1, 4.0, -0.15), (1.0, 0.3, 2.0, 0.0)]

RUN = r"""
import json, sys
sys.path.insert(0, {root!r})
from tpsc_conductivity_amd import capi
from tpsc_conductivity_amd.tpsc import CondOpt
pts = {pts!r}
g = CondOpt(32, 256, *pts[0][:2], params=(pts[0][3], 0.0, 0.0, pts[0][2]), coarse_pass=True)
out = []
for i, (n, T, U, tp) in enumerate(pts):
    if i:
        p = capi.Params(32, 256, n, T, 0.1, tp, 0.0, 0.0, U, 1, 0, 1.0)
        g.ctx.check(capi.lib().tpsc_set_params(g.ctx.handle, p), "set_params")
    g.solve()
    out.append({{k: g.scalar(k) for k in ("Usp", "Uch", "mu", "mu0", "kx0")}}
               | {{"jj": (g.get("chijj1") + g.get("chijj2")).tolist()}})
print(json.dumps(out))
"""


def run(no_graphs):
    env = dict(os.environ)
    env.pop("TPSC_NO_GRAPHS", None)
    env["TPSC_GRAPH_AFTER"] = "0"
    if no_graphs:
        env["TPSC_NO_GRAPHS"] = "1"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print(f"[graphs] 4 points, TPSC_NO_GRAPHS={int(no_graphs)}", flush=True)
    r = subprocess.run([sys.executable, "-c", RUN.format(root=root, pts=POINTS)], env=env, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_graph_replays_equal_eager_launches():
    a, b = run(False), run(True)
    for i, (x, y) in enumerate(zip(a, b)):
        for k in ("Usp", "Uch", "mu", "mu0", "kx0"):
            assert x[k] == y[k], (i, k)
        assert np.array_equal(np.array(x["jj"]), np.array(y["jj"])), i
