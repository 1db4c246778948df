"""Writes tests/golden/voxel_golden.json: the reference voxelizer's outputs on
the cases of voxel_cases.py.

Runs only in the development container (it reads /root/reference through
run_reference_voxelizer.js under Node); the GPU test machines never run it.  The JSON
holds data only: per case the mesh digest, the grid and options, sha256 of
the reference's tipo (uint8) and sdf (float32) arrays, its counts, and for the
small cases the full tipo as run-length pairs.

usage: python tests/golden/make_voxel_golden.py
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

from cfd__amd import meshes  # noqa: E402
from voxel_cases import cases  # noqa: E402


def main() -> None:
    out = {"generator": "tests/golden/make_voxel_golden.py",
           "reference": "src/geom/voxel.js voxelizar()", "cases": {}}
    with tempfile.TemporaryDirectory() as td:
        manifest = {"cases": []}
        meta = {}
        for name, make in cases().items():
            pos, idx, grid, banda, selar, with_rle = make()
            p, i = Path(td) / f"{name}.pos", Path(td) / f"{name}.idx"
            pos.tofile(p)
            idx.tofile(i)
            nx, ny, nz = grid
            manifest["cases"].append({"name": name, "pos": str(p), "idx": str(i),
                                      "grid": {"nx": nx, "ny": ny, "nz": nz},
                                      "banda": banda, "selar": selar, "with_rle": with_rle})
            meta[name] = {"mesh_sha256": meshes.mesh_digest(pos, idx), "grid": list(grid),
                          "banda": banda, "selar": selar, "n_tri": int(idx.size // 3)}
        mf = Path(td) / "manifest.json"
        mf.write_text(json.dumps(manifest))
        proc = subprocess.run(["node", "--experimental-vm-modules", "--no-warnings",
                               str(HERE / "run_reference_voxelizer.js"), str(mf)],
                              check=True, stdout=subprocess.PIPE, text=True)
        for line in proc.stdout.splitlines():
            r = json.loads(line)
            name = r.pop("name")
            meta[name].update(r)
        out["cases"] = meta
    lines = ["{", f' "generator": {json.dumps(out["generator"])},',
             f' "reference": {json.dumps(out["reference"])},', ' "cases": {']
    items = list(out["cases"].items())
    for k, (name, v) in enumerate(items):
        sep = "," if k < len(items) - 1 else ""
        lines.append(f"  {json.dumps(name)}: {json.dumps(v, separators=(',', ':'))}{sep}")
    lines += [" }", "}"]
    (HERE / "voxel_golden.json").write_text("\n".join(lines) + "\n")
    print(f"wrote {len(meta)} cases")


if __name__ == "__main__":
    main()
