"""Write the golden fixtures in tests/golden/ from the compiled QLens reference.

Runs only in a container that has /root/reference (to build oracle/_ref with oracle/ref.mk);
the fixtures it writes are plain data (inputs and the reference's outputs) so the tests run
anywhere.  Every input image is the reference's own mock of the workload (ref_harness mock:
plot_lensed_surface_brightness with simulated noise, lens.cpp:11882-12066), so the mock
generator, the PSF builder and the random generator are pinned by the same runs.  The
reference runs in a temporary directory that is removed afterwards.

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py C2 C4s     # only the named ones
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, REPO)

from oracle import oracle as O  # noqa: E402
from oracle import refrun  # noqa: E402
from qlens_amd import workloads as W  # noqa: E402

C4_SIDES = (20, 30, 40, 50, 60)


def _sub(work, name):
    d = os.path.join(work, name)
    os.makedirs(d, exist_ok=True)
    return d


def inversion_fixture(work, name, case, ray_stride=5, keep_model=True, psf_mode="input"):
    data = refrun.run_mock(case, _sub(work, name + "_mock"))
    sc, arr = refrun.run_invert(case, data, _sub(work, name), psf_mode=psf_mode)
    nx, ny = len(case["xvals"]) - 1, len(case["yvals"]) - 1
    # every ray_stride-th pixel's rays (in-mask pixels only carry reference rays)
    sel = np.arange(0, nx * ny, ray_stride)
    if case.get("mask") is not None:
        sel = sel[np.asarray(case["mask"]).reshape(-1)[sel] != 0]
    rays = arr["rays"].reshape(nx * ny, -1, 2)[sel]
    out = dict(data=data, minus2lnE=sc["minus2lnE"], chisq0=sc["chisq0"], logdetF=sc["logdetF"],
               logdetR=sc["logdetR"], n_img=int(sc["n_img_active"]), n_src=int(sc["n_src_active"]), src=arr["src"],
               mapped=arr["mapped"], active=arr["active"], ray_pixels=sel, rays=rays)
    if keep_model:
        out["model"] = arr["model"]
    if "psf" in arr:
        out["psf_used"] = arr["psf"]
    np.savez_compressed(os.path.join(HERE, "inversion_%s.npz" % name), **out)
    print(name, "-2lnE", repr(sc["minus2lnE"]), "n_img", sc["n_img_active"], "n_src", sc["n_src_active"], flush=True)


def c4s_fixture(work):
    """The 200 x 200 batched geometry on config 1's SIE + shear truth, one reference evaluation per
    source-grid side of the batched workload (20..60)."""
    base = W.c1_geometry_200(C4_SIDES[0])
    data = refrun.run_mock(base, _sub(work, "C4s_mock"))
    out = {"data": data, "sides": np.array(C4_SIDES)}
    nx, ny = len(base["xvals"]) - 1, len(base["yvals"]) - 1
    pix = np.arange(0, nx * ny, 7)
    out["model_pixels"] = pix
    for side in C4_SIDES:
        case = W.c1_geometry_200(side)
        sc, arr = refrun.run_invert(case, data, _sub(work, "C4s%d" % side))
        for k in ("minus2lnE", "chisq0", "logdetF", "logdetR"):
            out["%s_%d" % (k, side)] = sc[k]
        out["n_img_%d" % side] = int(sc["n_img_active"])
        out["n_src_%d" % side] = int(sc["n_src_active"])
        out["src_%d" % side] = arr["src"]
        out["model_%d" % side] = arr["model"].reshape(-1)[pix]
        print("C4s side", side, "-2lnE", repr(sc["minus2lnE"]), "n_src", sc["n_src_active"], flush=True)
    np.savez_compressed(os.path.join(HERE, "inversion_C4s.npz"), **out)


def deflection_fixture(work):
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
        # the shipped demo's lens (delaunay_fit_demo.in): alpha = 1.17, shear as components
        "sple_demo": [(W.SPLE, (1.3634, 1.17163, 0.963867, 81.9, 0.0102892, 0.00358392)),
                      (W.SHEAR_COMP, (0.0647257, -0.0575047, 0.0, 0.0))],
        "sple_steep": [(W.SPLE, (1.1, 1.6, 0.55, -30.0, 0.1, -0.2))],
        "sple_shallow": [(W.SPLE, (0.8, 0.45, 0.3, 65.0, -0.25, 0.15)), (W.SPLE, (0.5, 0.7, 1.0, 0.0, 0.3, 0.1))],
        "sple_alpha1": [(W.SPLE, (1.0, 1.0, 0.8, 10.0, 0.05, 0.0))],
        "sie_cored": [(W.SIE, (1.2, 0.7, 20.0, 0.05, -0.1, 0.15)), (W.SIE, (0.6, 1.0, 0.0, -0.4, 0.3, 0.2))],
    }
    out = {"points": pts}
    for name, lenses in lens_sets.items():
        r = refrun.run_deflect(lenses, pts, _sub(work, "defl_" + name), tol=1e-6)
        out[name + "_total"] = r["total"]
        out[name + "_per_lens"] = r["per_lens"]
        out[name + "_hess"] = r["hess"]
    np.savez_compressed(os.path.join(HERE, "deflection.npz"), **out)
    with open(os.path.join(HERE, "deflection_lenses.json"), "w") as f:
        json.dump({k: [[t, list(p)] for t, p in v] for k, v in lens_sets.items()}, f, indent=1)
    print("deflection fixture:", len(pts), "points x", len(lens_sets), "lens sets", flush=True)


def patterson_fixture(work):
    # digest of the reference's own tables (dumped by the harness), compared with the generated header
    import subprocess
    d = _sub(work, "pat")
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
    print("patterson digest written", flush=True)


def rng_fixture(work):
    """Normal deviates and raw draws of the reference generator (rand.h:19-76) for two seeds."""
    out = {}
    for seed, n in ((10, 4000), (987654321, 500)):
        dev, raw = refrun.run_rng(seed, n, _sub(work, "rng%d" % seed))
        out["deviates_%d" % seed] = dev
        out["int64_%d" % seed] = raw
    np.savez_compressed(os.path.join(HERE, "rng.npz"), **out)
    print("rng fixture written", flush=True)


PSF_CASES = [  # (sigma_x, sigma_y, threshold, xstep, ystep)
    (0.1 * W.FWHM_TO_SIGMA, 0.1 * W.FWHM_TO_SIGMA, 0.02, 0.05, 0.05),  # config 1: 5 x 5
    (0.1 * W.FWHM_TO_SIGMA, 0.1 * W.FWHM_TO_SIGMA, 0.01, 0.05, 0.05),  # the default threshold: 7 x 7
    (0.06, 0.035, 0.015, 0.05, 0.045),                                  # anisotropic (the CNS workload)
    (0.03, 0.09, 0.005, 0.04, 0.05),                                    # taller than wide
]


def psf_fixture(work):
    out = {"cases": np.array(PSF_CASES)}
    for k, (sx, sy, th, dx, dy) in enumerate(PSF_CASES):
        out["psf_%d" % k] = refrun.run_psf(sx, sy, th, dx, dy, _sub(work, "psf%d" % k))
    np.savez_compressed(os.path.join(HERE, "psf_gaussian.npz"), **out)
    print("psf fixture written", flush=True)


def cns_case():
    """The CNS workload with the Gaussian PSF the reference builds for its image grid:
    generate_PSF_matrix on the grid's own pixel steps ((xmax - xmin) / nx, as the inversion computes
    them), injected as the PSF matrix.  (Left to build it from psf_width itself, the reference's
    first inversion convolves with the previous PSF size: PSF_convolution_Lmatrix_dense reads
    psf_npixels_x/y before generate_PSF_matrix sets them, pixelgrid.cpp:13954-13955, 14097.)"""
    case = W.c_ns()
    sx, sy = case["psf_sigma"]
    xv, yv = case["xvals"], case["yvals"]
    dx, dy = (xv[-1] - xv[0]) / (len(xv) - 1), (yv[-1] - yv[0]) / (len(yv) - 1)
    with tempfile.TemporaryDirectory() as d:
        case["psf"] = refrun.run_psf(sx, sy, case["psf_threshold"], dx, dy, d)
    return case


def regopt_fixture(work):
    """'optimize_regparam on' (Brent over log10 lambda in [-3, 5], tol 0.01, at most 20 steps) on
    the C1 and CNS workloads, each on its own reference mock."""
    out = {}
    for name, case in (("C1", W.c1()), ("CNS", cns_case())):
        data = refrun.run_mock(case, _sub(work, "regopt_mock_" + name))
        case = dict(case, optimize_regparam=True)
        sc, arr = refrun.run_invert(case, data, _sub(work, "regopt_" + name))
        for k in ("minus2lnE", "chisq0", "logdetF", "logdetR", "regparam"):
            out["%s_%s" % (k, name)] = sc[k]
        out["n_src_%s" % name] = int(sc["n_src_active"])
        out["src_%s" % name] = arr["src"]
        out["data_%s" % name] = data
        print("regopt", name, "regparam", repr(sc["regparam"]), "-2lnE", repr(sc["minus2lnE"]), flush=True)
    np.savez_compressed(os.path.join(HERE, "regopt.npz"), **out)


AUTOGRID_CASES = [  # (name, workload, auto_srcgrid, auto_src_npixels)
    ("C1_both", "C1", True, True),
    ("CNS_box", "CNS", True, False),
    ("CNS_npix", "CNS", False, True),
    ("CNS_both", "CNS", True, True),
]


def autogrid_fixture(work):
    """'auto_srcgrid on' / 'auto_src_npixels on' (find_optimal_sourcegrid and
    find_optimal_sourcegrid_npixels, lens.cpp:12442-12458) on the C1 and CNS workloads, each on the
    data of its inversion fixture (the reference's mock of the same workload)."""
    out = {}
    cases = {"C1": W.c1(), "CNS": cns_case()}
    for tag, wl, ab, an in AUTOGRID_CASES:
        data = np.load(os.path.join(HERE, "inversion_%s.npz" % wl))["data"]
        case = dict(cases[wl], auto_srcgrid=ab, auto_src_npixels=an)
        sc, arr = refrun.run_invert(case, data, _sub(work, "auto_" + tag))
        for k in ("minus2lnE", "chisq0", "logdetF", "logdetR"):
            out["%s_%s" % (k, tag)] = sc[k]
        out["n_img_%s" % tag] = int(sc["n_img_active"])
        out["n_src_%s" % tag] = int(sc["n_src_active"])
        out["size_%s" % tag] = np.array([int(sc["src_nx"]), int(sc["src_ny"])])
        out["box_%s" % tag] = np.array([sc["src_xmin"], sc["src_xmax"], sc["src_ymin"], sc["src_ymax"]])
        out["src_%s" % tag] = arr["src"]
        out["active_%s" % tag] = arr["active"]
        print("autogrid", tag, "grid %dx%d" % (sc["src_nx"], sc["src_ny"]), "-2lnE", repr(sc["minus2lnE"]), flush=True)
    np.savez_compressed(os.path.join(HERE, "autogrid.npz"), **out)


def demo_fixture(work):
    """The shipped demo (workloads.demo) evaluated by the reference at the true lens and at the
    demo's starting guess.  The FITS inputs are read with the oracle's reader (the reference's
    cfitsio is absent here) and handed to the reference as its text image, mask and PSF."""
    case = W.demo()
    fits = os.path.join(HERE, "fits")
    img, _ = O.read_fits_np(os.path.join(fits, case["files"][0]))
    m, _ = O.read_fits_np(os.path.join(fits, case["files"][1]))
    p, _ = O.read_fits_np(os.path.join(fits, case["files"][2]))
    xv, yv = O.centred_grid(img.shape[0], img.shape[1], case["pixel_size"])
    case.update(xvals=xv, yvals=yv, psf=O.psf_crop(p, case["psf_threshold"]), mask=(m != 0).astype(np.uint8))
    out = {}
    for tag, lenses in (("true", W.DEMO_TRUE), ("start", W.DEMO_START)):
        sc, arr = refrun.run_invert(case, img, _sub(work, "demo_" + tag), lenses=lenses)
        for k in ("minus2lnE", "chisq0", "logdetF", "logdetR", "regparam"):
            out["%s_%s" % (k, tag)] = sc[k]
        out["n_img_%s" % tag] = int(sc["n_img_active"])
        out["n_src_%s" % tag] = int(sc["n_src_active"])
        out["size_%s" % tag] = np.array([int(sc["src_nx"]), int(sc["src_ny"])])
        out["box_%s" % tag] = np.array([sc["src_xmin"], sc["src_xmax"], sc["src_ymin"], sc["src_ymax"]])
        out["src_%s" % tag] = arr["src"]
        out["active_%s" % tag] = arr["active"]
        out["model_%s" % tag] = arr["model"]
        print("demo", tag, "grid %dx%d" % (sc["src_nx"], sc["src_ny"]), "regparam", repr(sc["regparam"]),
              "-2lnE", repr(sc["minus2lnE"]), "chisq0", sc["chisq0"], "n_img", sc["n_img_active"], flush=True)
    np.savez_compressed(os.path.join(HERE, "demo.npz"), **out)


FIXTURES = {
    "C1": lambda w: inversion_fixture(w, "C1", W.c1()),
    "C2s": lambda w: inversion_fixture(w, "C2s", W.c2_small()),
    "CNS": lambda w: inversion_fixture(w, "CNS", cns_case(), ray_stride=3),
    "C2": lambda w: inversion_fixture(w, "C2", W.c2(), ray_stride=9),
    "C4s": c4s_fixture,
    "CSPL": lambda w: inversion_fixture(w, "CSPL", W.c_sple(), ray_stride=2),
    "CSPS": lambda w: inversion_fixture(w, "CSPS", W.c_sple_steep(), ray_stride=3),
    "deflection": deflection_fixture,
    "patterson": patterson_fixture,
    "rng": rng_fixture,
    "psf": psf_fixture,
    "regopt": regopt_fixture,
    "autogrid": autogrid_fixture,
    "demo": demo_fixture,
}

if __name__ == "__main__":
    if not refrun.have_reference():
        sys.exit("oracle/_ref/ref_harness is not built (make -f oracle/ref.mk)")
    names = sys.argv[1:] or list(FIXTURES)
    with tempfile.TemporaryDirectory(prefix="qlx_golden_") as work:
        for n in names:
            FIXTURES[n](work)
