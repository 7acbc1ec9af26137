"""Object lifetimes across the C ABI: objects made on a context keep it
alive, so contexts and objects may be destroyed in any order. A garbage
collector finalizing module-level objects at interpreter exit runs their
finalizers in no fixed order (the context's first, here), which earlier
freed the context's stream under the population still to be destroyed."""
import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import pytest

from dnab import GpuContext, _native as N
from dnab import workloads as W

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent


def test_objects_outlive_their_context():
    c = GpuContext.try_init(0)
    assert c is not None
    w = W.config1(n_cand=8)
    ags, pop = W.device_objects(c, w)
    pop.evaluate()
    e, _ = pop.energies()
    N.lib.dnab_ctx_destroy(c.handle)      # the caller's hold goes first
    c._h = C.c_void_p()
    pop.close()                           # then the objects, which still sync on its stream
    for a in ags:
        a.close()
    c2 = GpuContext.try_init(0)           # the device is still usable
    ags2, pop2 = W.device_objects(c2, w)
    pop2.evaluate()
    e2, _ = pop2.energies()
    assert (e2 == e).all()
    c2.close()


SCRIPT = """
import sys
sys.path[:0] = [{pkg!r}, {root!r}]
import dnab
from dnab import workloads as W

def helper():          # a module-level function: the module dict becomes cyclic garbage at exit
    return 0

ctx = dnab.GpuContext.try_init(0)
w = W.config1(n_cand=8)
ags, pop = W.device_objects(ctx, w)
pop.evaluate()
ctx.set_timing(True)
pop.evaluate()
print("evaluated", flush=True)
"""


def test_module_level_objects_at_interpreter_exit():
    code = SCRIPT.format(pkg=str(ROOT / "de-novo-antigen-binder_amd"), root=str(ROOT))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ))
    assert "evaluated" in r.stdout, r.stdout + r.stderr
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
