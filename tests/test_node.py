"""The Node.js boundary: N-API addon over the C ABI plus SolverMI355X."""
import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cfd__amd import tunel

ROOT = Path(__file__).resolve().parent.parent
ADDON = ROOT / "cfd__amd" / "node" / "cfdmi.node"
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None or not ADDON.exists(),
                                reason="node or the built addon is missing")

EXPORTS = ["create", "destroy", "configure", "setTypes", "initEquilibrium", "setPopulations",
           "getPopulations", "setFront", "getFront", "step", "stepTimed", "forces",
           "updateMacros", "readMacros", "mass", "synchronize", "bytes", "voxelize",
           "setTuning", "version"]


def _node(code: str) -> str:
    return subprocess.run([NODE, "-e", code], cwd=ROOT, check=True, text=True,
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120).stdout


def test_addon_loads_and_exports():
    out = _node("const m=require('./cfd__amd/node/solver.js');"
                "console.log(JSON.stringify(Object.keys(m.addon)));"
                "console.log(typeof m.SolverMI355X.criar, typeof m.voxelizar)")
    keys = json.loads(out.splitlines()[0])
    assert sorted(keys) == sorted(EXPORTS)
    assert out.splitlines()[1] == "function function"


def test_errors_become_js_errors():
    """Argument errors of the C ABI surface as thrown Errors with its text."""
    out = _node("const {addon}=require('./cfd__amd/node/solver.js');"
                "try { addon.create(0, 4, 4, false); console.log('no throw'); }"
                "catch (e) { console.log(e instanceof Error, e.message); }")
    assert out.strip().startswith("true cfd_create: grid 0x4x4 must be positive")


@pytest.mark.gpu
def test_node_run_matches_python_run(tmp_path):
    """Same case through the JS class and through the Python mirror: the
    macro field and the forces are bit-identical (one library, one kernel)."""
    from cfd__amd.solver import SolverMI355X
    nx, ny, nz = 40, 20, 16
    g = {"nx": nx, "ny": ny, "nz": nz}
    corpo = tunel.corpoAnalitico(g, {"tipo": "esfera", "cx": 14, "cy": 10, "cz": 8, "raio": 4})
    tipo, _, _ = tunel.montarTipos(g, corpo, esteira=True)
    tipo.astype(np.uint32).tofile(tmp_path / "tipo.u32")
    subprocess.run([NODE, str(ROOT / "tests" / "node_case.js"), str(tmp_path / "tipo.u32"),
                    str(nx), str(ny), str(nz), str(tmp_path / "js")], check=True, timeout=120)
    js = json.loads((tmp_path / "js.json").read_text())
    mjs = np.fromfile(tmp_path / "js.macros", dtype=np.float32)

    esponja = {"inicio": nx - 1 - 8, "comprimento": 8}
    s = SolverMI355X.criar(g)
    try:
        s.definirTipos(tipo)
        s.configurar(omegaPlus=1.6, lesCs=0.1, inletU=[0.05, 0, 0], beltU=[0.05, 0, 0],
                     esponja=esponja)
        off = s.calibrarRepouso()
        s.rampa(40, lambda f: s.configurar(omegaPlus=1.6, lesCs=0.1, inletU=[0.05 * f, 0, 0],
                                            beltU=[0.05 * f, 0, 0], esponja=esponja))
        s.inicializar()
        for _ in range(6):
            s.passo(10)
        f = s.medirForcas()
        m = s.lerMacros()
    finally:
        s.destruir()
    assert js["passos"] == 60 and js["frente"] == "B"
    assert np.allclose(js["off"], off, rtol=0, atol=0)
    assert np.array_equal(np.array(js["f"]), np.array(f))
    assert np.array_equal(mjs, m)
