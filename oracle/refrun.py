"""Run the compiled QLens reference (oracle/_ref/ref_harness) on a workload -- checker only.

Writes the workload's image data as QLens text files (root.dat x-major, root.x, root.y, all
with 17 significant digits so every double round-trips), a QLens command script, and runs the
harness.  Used by tests/golden/make_golden.py (fixtures) and by bench.py's cpu_baseline leg;
never by the product.
"""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, ".."))
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")

SIE, NFW, SHEAR, SPLE, SHEAR_COMP = 1, 2, 3, 4, 5


def have_reference():
    return os.path.exists(HARNESS)


def _fmt(v):
    return "%.17g" % v


def lens_commands(lenses):
    out = []
    comps = any(t == SHEAR_COMP for t, _ in lenses)
    if comps:
        if any(t == SHEAR for t, _ in lenses):
            raise ValueError("the reference takes shear as components or as (gamma, theta), not both")
        out.append("shear_components on")
    for t, p in lenses:
        if t == SIE:  # lens sple <b> <alpha> <s> <q> [theta] [xc] [yc]  (commands.cpp:465)
            b, q, th, xc, yc, s = p
            out.append("lens sple %s 1 %s %s %s %s %s" % (_fmt(b), _fmt(s), _fmt(q), _fmt(th), _fmt(xc), _fmt(yc)))
        elif t == NFW:  # lens nfw <ks> <rs> <q> [theta] [xc] [yc]  (commands.cpp:549)
            ks, rs, q, th, xc, yc = p
            out.append("lens nfw %s %s %s %s %s %s" % (_fmt(ks), _fmt(rs), _fmt(q), _fmt(th), _fmt(xc), _fmt(yc)))
        elif t == SHEAR:  # lens shear <shear> [theta] [xc] [yc]  (commands.cpp:492)
            g, th, xc, yc = p
            out.append("lens shear %s %s %s %s" % (_fmt(g), _fmt(th), _fmt(xc), _fmt(yc)))
        elif t == SHEAR_COMP:  # with shear_components on: lens shear <shear1> <shear2> [xc] [yc]
            g1, g2, xc, yc = p
            out.append("lens shear %s %s %s %s" % (_fmt(g1), _fmt(g2), _fmt(xc), _fmt(yc)))
        elif t == SPLE:  # lens sple <b> <alpha> <s> <q> [theta] [xc] [yc], s = 0
            b, al, q, th, xc, yc = p
            out.append("lens sple %s %s 0 %s %s %s %s" % (_fmt(b), _fmt(al), _fmt(q), _fmt(th), _fmt(xc), _fmt(yc)))
    return out


def write_inputs(case, data, workdir, lenses=None, psf_mode="input"):
    os.makedirs(workdir, exist_ok=True)
    xv, yv = case["xvals"], case["yvals"]
    with open(os.path.join(workdir, "img.x"), "w") as f:
        f.write("\n".join(_fmt(v) for v in xv) + "\n")
    with open(os.path.join(workdir, "img.y"), "w") as f:
        f.write("\n".join(_fmt(v) for v in yv) + "\n")
    with open(os.path.join(workdir, "img.dat"), "w") as f:
        for row in np.asarray(data):
            f.write(" ".join(_fmt(v) for v in row) + "\n")
    psf = np.ascontiguousarray(case["psf"], dtype=np.float64)
    psf.tofile(os.path.join(workdir, "psf.bin"))
    lines = [
        "fit source_mode cartesian",
        "split_imgpixels %s" % ("on" if case["nsplit"] else "off"),
        "imgpixel_nsplit %d" % max(case["nsplit"], 1),
        "raytrace_method interpolate_3pt",
        "inversion_method dense",
        "fit regularization curvature",
        "integral_tolerance %s" % _fmt(case["integral_tol"]),
        "srcgrid %s %s %s %s" % tuple(_fmt(v) for v in case["srcgrid"]),
        "src_npixels %d %d" % tuple(case["src_n"]),
        "auto_srcgrid %s" % ("on" if case.get("auto_srcgrid") else "off"),
        "auto_src_npixels %s" % ("on" if case.get("auto_src_npixels") else "off"),
        "regparam %s" % _fmt(case["regparam"]),
        "optimize_regparam %s" % ("on" if case.get("optimize_regparam") else "off"),
        "regparam_minlog %s" % _fmt(case.get("regopt_bounds", (-3.0, 5.0))[0]),
        "regparam_maxlog %s" % _fmt(case.get("regopt_bounds", (-3.0, 5.0))[1]),
        "bg_pixel_noise %s" % _fmt(case["sigma"]),
    ]
    if psf_mode == "gaussian":
        sig = case["psf_sigma"]
        sig = (sig, sig) if np.isscalar(sig) else tuple(sig)
        lines += ["psf_width %s %s" % (_fmt(sig[0]), _fmt(sig[1])), "psf_threshold %s" % _fmt(case["psf_threshold"])]
    lines += lens_commands(case["lenses"] if lenses is None else lenses)
    lines += ["sbmap loadimg %s" % os.path.join(workdir, "img")]
    with open(os.path.join(workdir, "script.in"), "w") as f:
        f.write("\n".join(lines) + "\n")
    return os.path.join(workdir, "script.in")


def read_scalars(outdir):
    out = {}
    with open(os.path.join(outdir, "scalars.txt")) as f:
        for line in f:
            k, v = line.split()
            out[k] = float(v)
    return out


def run_invert(case, data, workdir, lenses=None, nrep=1, psf_mode="input", threads=None):
    script = write_inputs(case, data, workdir, lenses=lenses, psf_mode=psf_mode)
    args = [HARNESS, "invert", script, workdir, _fmt(case["sigma"]), str(nrep)]
    if psf_mode == "input":
        psf = np.asarray(case["psf"])
        args.append("psf=%s:%d:%d" % (os.path.join(workdir, "psf.bin"), psf.shape[0], psf.shape[1]))
    if case.get("mask") is not None:
        mk = np.ascontiguousarray(case["mask"], dtype=np.uint8)
        mk.tofile(os.path.join(workdir, "mask.bin"))
        args.append("mask=%s:%d:%d" % (os.path.join(workdir, "mask.bin"), mk.shape[0], mk.shape[1]))
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    r = subprocess.run(args, cwd=workdir, capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError("reference harness failed (%d): %s %s" % (r.returncode, r.stdout[-2000:], r.stderr[-2000:]))
    sc = read_scalars(workdir)
    nx, ny = len(case["xvals"]) - 1, len(case["yvals"]) - 1
    ns = max(case["nsplit"], 1)
    snx, sny = int(sc["src_nx"]), int(sc["src_ny"])  # the automatic size, when asked for
    arr = dict(
        rays=np.fromfile(os.path.join(workdir, "rays.bin")).reshape(nx, ny, ns * ns, 2),
        mapped=np.fromfile(os.path.join(workdir, "mapped.bin"), dtype=np.uint8).reshape(nx, ny),
        model=np.fromfile(os.path.join(workdir, "model.bin")).reshape(nx, ny),
        src=np.fromfile(os.path.join(workdir, "src.bin")).reshape(snx, sny),
        active=np.fromfile(os.path.join(workdir, "active.bin"), dtype=np.int32).reshape(snx, sny),
    )
    if os.path.exists(os.path.join(workdir, "psf_used.bin")):
        arr["psf"] = np.fromfile(os.path.join(workdir, "psf_used.bin")).reshape(int(sc["psf_nx"]), int(sc["psf_ny"]))
    return sc, arr


def _run(args, workdir, threads=None):
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    r = subprocess.run(args, cwd=workdir, capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError("reference harness failed (%d): %s %s" % (r.returncode, r.stdout[-2000:], r.stderr[-2000:]))
    return r


def run_mock(case, workdir, sigma=None, seed=None):
    """The reference's own mock image of a workload (ref_harness mock): the case's lenses and
    Sersic sources on the no-data image grid spanned by its corners, its PSF matrix injected, and
    pixel noise sigma from the reference generator seeded with seed (sigma 0: no noise)."""
    os.makedirs(workdir, exist_ok=True)
    xv, yv = case["xvals"], case["yvals"]
    nx, ny = len(xv) - 1, len(yv) - 1
    sigma = case["sigma"] if sigma is None else sigma
    seed = case["seed"] if seed is None else seed
    lines = [
        "fit source_mode sbprofile",
        "split_imgpixels %s" % ("on" if case["nsplit"] else "off"),
        "imgpixel_nsplit %d" % max(case["nsplit"], 1),
        "grid %s %s %s %s" % (_fmt(xv[0]), _fmt(xv[-1]), _fmt(yv[0]), _fmt(yv[-1])),
        "img_npixels %d %d" % (nx, ny),
        "integral_tolerance %s" % _fmt(case["integral_tol"]),
    ]
    lines += lens_commands(case["lenses"])
    for (s0, reff, n, q, th, xc, yc) in case["sources"]:
        lines.append("source sersic %s %s %s %s %s %s %s" % tuple(_fmt(v) for v in (s0, reff, n, q, th, xc, yc)))
    if sigma > 0:
        lines += ["bg_pixel_noise %s" % _fmt(sigma), "simulate_pixel_noise on", "random_seed %d" % seed]
    script = os.path.join(workdir, "script.in")
    with open(script, "w") as f:
        f.write("\n".join(lines) + "\n")
    psf = np.ascontiguousarray(case["psf"], dtype=np.float64)
    psf.tofile(os.path.join(workdir, "psf.bin"))
    _run([HARNESS, "mock", script, workdir, "psf=%s:%d:%d" % (os.path.join(workdir, "psf.bin"), psf.shape[0], psf.shape[1])],
         workdir)
    return np.fromfile(os.path.join(workdir, "mock.bin")).reshape(nx, ny)


def run_rng(seed, n, workdir):
    """First n normal deviates and n raw 64-bit draws of the reference generator (rand.h)."""
    os.makedirs(workdir, exist_ok=True)
    script = os.path.join(workdir, "script.in")
    with open(script, "w") as f:
        f.write("\n")
    _run([HARNESS, "rng", script, workdir, str(int(seed)), str(int(n))], workdir)
    return (np.fromfile(os.path.join(workdir, "deviates.bin")),
            np.fromfile(os.path.join(workdir, "int64.bin"), dtype=np.uint64))


def run_psf(sigma_x, sigma_y, threshold, xstep, ystep, workdir):
    """QLens::generate_PSF_matrix for psf_width sigma_x sigma_y and psf_threshold on xstep x ystep pixels."""
    os.makedirs(workdir, exist_ok=True)
    script = os.path.join(workdir, "script.in")
    with open(script, "w") as f:
        f.write("psf_width %s %s\npsf_threshold %s\n" % (_fmt(sigma_x), _fmt(sigma_y), _fmt(threshold)))
    _run([HARNESS, "psf", script, workdir, _fmt(xstep), _fmt(ystep)], workdir)
    sc = read_scalars(workdir)
    return np.fromfile(os.path.join(workdir, "psf.bin")).reshape(int(sc["psf_nx"]), int(sc["psf_ny"]))


def run_deflect(lenses, xy, workdir, tol=1e-6):
    os.makedirs(workdir, exist_ok=True)
    lines = ["integral_tolerance %s" % _fmt(tol)] + lens_commands(lenses)
    script = os.path.join(workdir, "script.in")
    with open(script, "w") as f:
        f.write("\n".join(lines) + "\n")
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
    pts = os.path.join(workdir, "points.bin")
    xy.tofile(pts)
    r = subprocess.run([HARNESS, "deflect", script, workdir, "points=%s:%d" % (pts, len(xy))], cwd=workdir,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("reference harness failed: %s %s" % (r.stdout[-2000:], r.stderr[-2000:]))
    n, nl = len(xy), len(lenses)
    return dict(
        total=np.fromfile(os.path.join(workdir, "def_total.bin")).reshape(n, 2),
        per_lens=np.fromfile(os.path.join(workdir, "def_per_lens.bin")).reshape(nl, n, 2),
        hess=np.fromfile(os.path.join(workdir, "hess_per_lens.bin")).reshape(nl, n, 2, 2),
    )
