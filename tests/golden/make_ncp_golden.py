"""Extract the NCP-level fixtures the sensitivity tests use, from the reference's data files.

Reads (as data: JSON and CSV) data/ncp/scenes/NN.json of the reference's contact-scene-v1 package and
data/ncp/adjoint_fix_table.csv, and writes
  ncp_scenes.npz          G, b, mu of the listed body-contact scenes, assembled as the package's
                          schema states: J = per-contact blocks [D ; r x D] (rows n, t1, t2; columns
                          v, omega of body a, minus the same for body b), M^-1 block diagonal of the
                          bodies' inverse masses and inertias, G = J M^-1 J^T + diag(eta)
  ncp_adjoint_table.json  the all-slip rows of the adjoint table (labels, natural residual and
                          condition number of the active-set Jacobian)
Run from the repository root with the reference tree mounted:

    python tests/golden/make_ncp_golden.py /root/reference
"""
import csv
import json
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent
SCENES = (0,)  # the hyperstatic 4-point cube: a singular active-set Jacobian


def assemble(sc):
    nc, bodies, contacts = int(sc["n_c"]), sc["bodies"], sc["contacts"]
    dof = int(bodies[0]["dof"])
    nv = dof * len(bodies)
    Minv = np.zeros((nv, nv))
    for bd in bodies:
        k = int(bd["id"])
        im = float(bd["inv_mass"])
        if dof == 3:
            Minv[3 * k:3 * k + 3, 3 * k:3 * k + 3] = im * np.eye(3)
        else:
            Minv[6 * k:6 * k + 3, 6 * k:6 * k + 3] = im * np.eye(3)
            Minv[6 * k + 3:6 * k + 6, 6 * k + 3:6 * k + 6] = np.diag(np.asarray(bd["inv_inertia"], float))
    J = np.zeros((3 * nc, nv))
    for ct in contacts:
        c = int(ct["id"])
        D = np.array([ct["normal"], ct["t1"], ct["t2"]], float)
        for body, lever, sign in ((int(ct["a"]), ct.get("lever_a"), 1.0), (int(ct["b"]), ct.get("lever_b"), -1.0)):
            if body < 0:
                continue
            J[3 * c:3 * c + 3, dof * body:dof * body + 3] += sign * D
            if dof == 6:
                J[3 * c:3 * c + 3, dof * body + 3:dof * body + 6] += sign * np.cross(np.asarray(lever, float)[None, :], D)
    G = J @ Minv @ J.T
    eta = np.asarray(sc["eta"], float)
    if np.any(eta):
        G = G + np.diag(np.repeat(eta, 3))
    return G, np.asarray(sc["b"], float), np.asarray(sc["mu"], float)


def main(ref_root):
    root = Path(ref_root) / "data" / "ncp"
    arrays, names = {}, {}
    for i in SCENES:
        sc = json.loads((root / "scenes" / f"{i:02d}.json").read_text())
        if sc["model"] != "body_contacts":
            continue
        G, b, mu = assemble(sc)
        arrays[f"s{i:02d}_G"], arrays[f"s{i:02d}_b"], arrays[f"s{i:02d}_mu"] = G, b, mu
        names[f"s{i:02d}"] = sc["scene"]
    np.savez_compressed(OUT / "ncp_scenes.npz", **arrays)
    with (root / "adjoint_fix_table.csv").open() as fh:
        rows = [r for r in csv.DictReader(fh) if r["scene"].startswith("all_slip_")]
    table = {r["scene"]: {"n_c": int(r["n_c"]), "labels": r["labels"],
                          "natural_residual": float(r["natural_residual"]),
                          "cond_dFdlam": float(r["cond_dFdlam"]),
                          "fixed_max_rel_dlam_db": float(r["fixed_max_rel_dlam_db"]),
                          "fixed_max_rel_dlam_dmu": float(r["fixed_max_rel_dlam_dmu"])} for r in rows}
    meta = {"source": "data/ncp/scenes/*.json, data/ncp/adjoint_fix_table.csv", "scenes": names,
            "all_slip": table}
    (OUT / "ncp_adjoint_table.json").write_text(json.dumps(meta, indent=1) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
