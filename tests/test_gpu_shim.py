"""The C++ drop-in (include/loadsg_amd_shim.hpp) against the compiled reference in one
program (oracle/shim_check.cpp): reference MatrixVectorHomogen::multiplication and
CG::solveHomogen on the CPU vs the same calls through the drop-in on the GPU, on the
same AdaptiveSparseGrid and VectorSparseG objects."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


def _run(d, level):
    exe = os.path.join(REF, "shim_check_d%d" % d)
    if not os.path.exists(exe):
        pytest.skip("%s not built (make -C oracle shim needs the reference sources)" % exe)
    print("shim_check d=%d level=%d" % (d, level), flush=True)
    r = subprocess.run([exe, str(level)], capture_output=True, text=True, timeout=300)
    line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
    return r.returncode, json.loads(line)


def test_shim_links_cpu():
    """Without a GPU the drop-in loads the library and reports that no device is present."""
    exe = os.path.join(REF, "shim_check_d3")
    if not os.path.exists(exe):
        pytest.skip("shim_check_d3 not built")
    rc, res = _run(3, 3)
    assert rc in (0, 2), res
    if rc == 2:
        assert res["error"] == "no GPU"


@pytest.mark.gpu
@pytest.mark.parametrize("d,level", [(3, 6), (5, 4)])
def test_shim_vs_reference(d, level):
    rc, res = _run(d, level)
    print(json.dumps(res), flush=True)
    assert rc == 0, res
    assert res["matvec_poisson_rel"] < 1e-12 and res["matvec_helm_rel"] < 1e-12
    assert res["cg_iters_ref"] == res["cg_iters_dev"]
    assert res["cg_x_rel"] < 1e-10
