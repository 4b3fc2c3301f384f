"""Extract small golden fixtures from the reference's archived trajectories.

Reads (as data only, allow_pickle=False) the reference's cross-hardware archive
reports/position_solve_matrix/inputs/cross_hardware_l4.npz and its JSON receipt, and writes
the subsets the tests use. Run from the repository root with the reference tree mounted:

    python tests/golden/make_golden.py /root/reference
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent


def main(ref_root: str) -> None:
    base = Path(ref_root) / "reports/position_solve_matrix/inputs/cross_hardware_l4"
    receipt = json.loads(Path(str(base) + ".json").read_text())
    with np.load(str(base) + ".npz", allow_pickle=False) as d:
        arrays = {}
        meta = {"source": "reports/position_solve_matrix/inputs/cross_hardware_l4.npz",
                "solver_params": {"vit": 40, "pit": 10, "dims": [0.3, 0.3, 0.2], "mu": 0.5,
                                  "dt": 1 / 240, "substeps": 1, "density": 700.0},
                "records": {}}
        for rec in receipt["records"]:
            eng, scene = rec["engine"], rec["scene"]
            leg = rec["legs"][0]
            meta["records"][f"{eng}__{scene}"] = {
                "initial": leg["initial"], "contacts": leg["contacts"], "exact_repeat": rec["exact"]}
            pre = f"{eng}__{scene}__0__"
            if scene == "stack8":
                for k in ("xc", "vc", "om", "Rm", "force"):
                    arrays[f"{eng}__stack8__{k}"] = d[pre + k]
            else:
                steps = [0, 3, 4, 9, 19, 39]
                arrays[f"{eng}__lattice1000__steps"] = np.array(steps, np.int32)
                for k in ("xc", "vc", "om"):
                    arrays[f"{eng}__lattice1000__{k}"] = d[pre + k][steps]
                meta["records"][f"{eng}__lattice1000"]["sha256_per_step"] = {
                    k: [hashlib.sha256(np.ascontiguousarray(d[pre + k][s]).tobytes()).hexdigest()
                        for s in range(d[pre + k].shape[0])] for k in ("xc", "vc", "om")}
    np.savez_compressed(OUT / "reference_trajectories.npz", **arrays)
    (OUT / "reference_trajectories.json").write_text(json.dumps(meta, indent=1) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
