"""Handle teardown in any order.  A surrogate and a captured IRT plan each hold a share of
their context, so destroying the context first leaves them working, and destroying them
afterwards frees it.  The run happens in a child process, so that a use-after-free or a
crash at interpreter exit shows up as its exit status."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
from dna_amd import Context, Surrogate, default_params
ctx = Context(0)
p = default_params(1000.0, 24)
g = ctx.transport(p, copy=1)
n_therm = int((g["results"]["esc"] == 0).sum())
s = Surrogate(ctx)
X = np.linspace(-1, 1, 4 * 70, dtype=np.float32).reshape(4, 70)
y0 = s.forward(X)
plan = ctx.irt_plan(None, [0, g["rad_n"]], [n_therm], [1000.0], [1.0, 1000.0])
r0 = plan.run(5)[0]
ctx.close()                                  # the creator's handle first
r1 = plan.run(5)[0]                          # the plan still runs on its context
assert [list(r0.tally[k]) for k in range(2)] == [list(r1.tally[k]) for k in range(2)]
assert np.array_equal(s.forward(X), y0)      # and so does the surrogate
plan.close()
s.close()                                    # the last share: the context is freed here
print("teardown ok", flush=True)
'''


def test_context_destroyed_before_its_dependants():
    print("child process: context destroyed before its plan and surrogate", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, os.path.join(ROOT, "webgpu-dna_amd")],
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "teardown ok" in r.stdout
