"""Write the golden fixtures in tests/golden/ from the compiled QLens reference.

Runs only in a container that has /root/reference (to build oracle/_ref with oracle/ref.mk);
the fixtures it writes are plain data (inputs and the reference's outputs) so the tests can run
anywhere.  Inputs are seeded mock images made by the oracle (whose mock path is itself compared
with the reference's in the 'mock' fixture).

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

from oracle import oracle as O  # noqa: E402
from oracle import refrun  # noqa: E402
from qlens_amd import workloads as W  # noqa: E402

WORK = os.path.join(REPO, "_work", "golden_runs")


def inversion_fixture(name, case):
    data = O.mock_image(case)
    sc, arr = refrun.run_invert(case, data, os.path.join(WORK, name))
    nx, ny = len(case["xvals"]) - 1, len(case["yvals"]) - 1
    # keep every 5th pixel's rays to stay small; the full arrays are checked through the evidence
    sel = np.arange(0, nx * ny, 5)
    rays = arr["rays"].reshape(nx * ny, -1, 2)[sel]
    np.savez_compressed(
        os.path.join(HERE, "inversion_%s.npz" % name),
        data=data, minus2lnE=sc["minus2lnE"], chisq0=sc["chisq0"], logdetF=sc["logdetF"], logdetR=sc["logdetR"],
        n_img=int(sc["n_img_active"]), n_src=int(sc["n_src_active"]), src=arr["src"], model=arr["model"],
        mapped=arr["mapped"], active=arr["active"], ray_pixels=sel, rays=rays)
    print(name, "-2lnE", repr(sc["minus2lnE"]), "n_img", sc["n_img_active"], "n_src", sc["n_src_active"])


def deflection_fixture():
    rng = np.random.default_rng(7)
    pts = np.concatenate([
        rng.uniform(-3, 3, size=(200, 2)),
        rng.uniform(-0.05, 0.05, size=(20, 2)),                      # near the centre (small x^2 branch)
        np.array([[0.0, 0.0], [1e-15, 0.0], [5.0, 0.0], [0.0, 5.0],  # exact centre, |x^2-1| branch at r=rs
                  [4.99999, 0.0], [0.0, 3.7], [-2.0, 1.0]]),
    ])
    lens_sets = {
        "nfw_sie_shear": [(W.NFW, (0.2, 5.0, 0.75, -15.0, 0.0, 0.0)), (W.SIE, (1.0, 0.85, 0.0, 0.0, 0.0, 0.0)),
                          (W.SHEAR, (0.03, 0.0, 0.0, 0.0))],
        "sie_shear_c1": [(W.SIE, (1.2, 0.8, 30.0, 0.0, 0.0, 0.0)), (W.SHEAR, (0.05, 20.0, 0.0, 0.0))],
        "offcentre": [(W.NFW, (0.35, 1.5, 0.5, 70.0, 0.2, -0.1)), (W.SIE, (0.7, 0.6, -40.0, -0.3, 0.25, 0.05)),
                      (W.SHEAR, (0.08, 135.0, 0.1, 0.0))],
        "spherical": [(W.NFW, (0.3, 2.0, 1.0, 0.0, 0.0, 0.0)), (W.SIE, (0.9, 1.0, 0.0, 0.1, 0.0, 0.0))],
    }
    out = {"points": pts}
    for name, lenses in lens_sets.items():
        r = refrun.run_deflect(lenses, pts, os.path.join(WORK, "defl_" + name), tol=1e-6)
        out[name + "_total"] = r["total"]
        out[name + "_per_lens"] = r["per_lens"]
        out[name + "_hess"] = r["hess"]
    np.savez_compressed(os.path.join(HERE, "deflection.npz"), **out)
    import json
    with open(os.path.join(HERE, "deflection_lenses.json"), "w") as f:
        json.dump({k: [[t, list(p)] for t, p in v] for k, v in lens_sets.items()}, f, indent=1)
    print("deflection fixture:", len(pts), "points x", len(lens_sets), "lens sets")


def patterson_fixture():
    # digest of the reference's own tables (dumped by the harness), compared with the generated header
    import subprocess
    d = os.path.join(WORK, "pat")
    os.makedirs(d, exist_ok=True)
    script = os.path.join(d, "script.in")
    with open(script, "w") as f:
        f.write("lens nfw 0.2 5 0.75 0 0 0\n")
    subprocess.run([refrun.HARNESS, "patterson", script, d], check=True, capture_output=True)
    pts = np.fromfile(os.path.join(d, "pat_points.bin"))
    w = np.fromfile(os.path.join(d, "pat_weights.bin"))
    with open(os.path.join(HERE, "patterson_digest.txt"), "w") as f:
        f.write("points_sha256 %s\n" % hashlib.sha256(pts.tobytes()).hexdigest())
        f.write("weights_sha256 %s\n" % hashlib.sha256(w.tobytes()).hexdigest())
        f.write("n_points %d\nn_weights %d\n" % (len(pts), len(w)))
    print("patterson digest written")


if __name__ == "__main__":
    if not refrun.have_reference():
        sys.exit("oracle/_ref/ref_harness is not built (make -f oracle/ref.mk)")
    os.makedirs(WORK, exist_ok=True)
    inversion_fixture("C1", W.c1())
    inversion_fixture("C2s", W.c2_small())
    deflection_fixture()
    patterson_fixture()
