"""Writes tests/golden/prepare_golden.json: the reference's own ler() and
preparar() (src/geom/parse.js, prepare.js) on the mesh fixtures.

Runs only in the development container (it reads /root/reference through
run_reference_prepare.js under Node); the GPU test machines never run it.
The JSON holds data only: per case the file, the grid and the preparar
options, sha256 of the reference's read positions / indices and of its placed
positions (float32), and the placement's scalars.

usage: python tests/golden/make_prepare_golden.py
"""
from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(HERE))

from prepare_cases import cases  # noqa: E402


def main() -> None:
    out = {"generator": "tests/golden/make_prepare_golden.py",
           "reference": "src/geom/parse.js ler(), src/geom/prepare.js preparar()", "cases": {}}
    with tempfile.TemporaryDirectory() as td:
        manifest = {"cases": []}
        for c in cases(Path(td)):
            manifest["cases"].append({"id": c["id"], "file": str(c["file"]), "name": c["name"],
                                      "grade": c["grade"], "opcoes": c["opcoes"]})
            out["cases"][c["id"]] = {"name": c["name"], "grade": c["grade"],
                                     "opcoes": c["opcoes"]}
        mf = Path(td) / "manifest.json"
        mf.write_text(json.dumps(manifest))
        res = subprocess.run(["node", "--experimental-vm-modules", "--harmony-nullish",
                              "--harmony-optional-chaining",
                              str(HERE / "run_reference_prepare.js"), str(mf)],
                             check=True, text=True, stdout=subprocess.PIPE,
                             stderr=subprocess.PIPE).stdout
    for line in res.splitlines():
        r = json.loads(line)
        out["cases"][r.pop("case")].update(r)
    (HERE / "prepare_golden.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"{len(out['cases'])} cases -> {HERE / 'prepare_golden.json'}")


if __name__ == "__main__":
    main()
