"""Concrete, seeded workloads for the pixellated-source evidence (shared by tests and bench).

Each builder returns a dict describing one problem in QLens conventions:
  xvals, yvals   image pixel corners (nx+1, ny+1); x_i = xmin + i*((xmax-xmin)/nx), the same
                 multiply-based corners the reference's no-data image grid uses
                 (pixelgrid.cpp:9582-9595)
  nsplit         subpixel split per axis (QLens default_imgpixel_nsplit, lens.cpp:497)
  lenses         [(type, params)]: SIE (b, q, theta, xc, yc, s), NFW (ks, rs, q, theta, xc, yc),
                 SHEAR (gamma, theta, xc, yc); lens angles from +y (profile.cpp:20)
  sources        Sersic (s0, Reff, n, q, theta, xc, yc) used only to simulate the data
  psf            PSF matrix psf[i][j] (QLens psf_matrix orientation)
  sigma, seed    pixel noise and the NR3 generator seed (rand.h:19-76)
  srcgrid, src_n Cartesian source grid bounds and cell counts (auto_srcgrid / auto_src_npixels off)
  regparam       curvature regularization lambda
  integral_tol   Gauss-Patterson relative tolerance (integral_tolerance)
Sizes and parameter values follow the five workloads of the benchmark description:
  C1 SIE+shear 100^2 image / 30^2 source, Gaussian PSF, FP64
  C2 NFW(quadrature)+SIE+shear 200^2 / 60^2, Moffat 21x21 PSF
  C3 surrogate deflection evaluation on a 400^2 x 4x4 image
  C4 256 proposals around C1's truth on a 200^2 image with uneven source grids (c4; c4_nfw is
     the same batch shape on C2's numerical-profile truth)
  C5 400^2 image / 100^2 source, mixed precision, batches of 16
  CNS a non-square, masked, anisotropic-PSF orientation check (c_ns)
"""
import numpy as np

SIE, NFW, SHEAR, SPLE, SHEAR_COMP = 1, 2, 3, 4, 5
FWHM_TO_SIGMA = 1.0 / 2.35482


def corners(xmin, xmax, n):
    step = (xmax - xmin) / n
    return np.array([xmin + i * step for i in range(n + 1)], dtype=np.float64)


def gaussian_psf(sigma, threshold, xstep, ystep):
    """QLens::generate_PSF_matrix without supersampling (pixelgrid.cpp:14601-14668), restated
    in numpy for workload setup: point-sampled Gaussian truncated at psf_threshold."""
    frac = np.sqrt(-2 * np.log(threshold))
    nxd, nyd = frac * sigma / xstep, frac * sigma / ystep
    nxh, nyh = int(nxd), int(nyd)
    if nxd - nxh > 0.5:
        nxh += 1
    if nyd - nyh > 0.5:
        nyh += 1
    nx, ny = 2 * nxh + 1, 2 * nyh + 1
    xmax = ((2 * nxh + 1) * xstep - xstep) / 2
    ymax = ((2 * nyh + 1) * ystep - ystep) / 2
    psf = np.zeros((nx, ny))
    x = -xmax
    for i in range(nx):
        y = -ymax
        for j in range(ny):
            psf[i, j] = np.exp(-0.5 * ((x / sigma) ** 2 + (y / sigma) ** 2))
            y += ystep
        x += xstep
    return psf / psf.sum()


def gaussian_psf_xy(sigma_x, sigma_y, threshold, xstep, ystep):
    """gaussian_psf with separate widths along x and y (QLens 'psf_width <x> <y>')."""
    frac = np.sqrt(-2 * np.log(threshold))
    nxd, nyd = frac * sigma_x / xstep, frac * sigma_y / ystep
    nxh, nyh = int(nxd), int(nyd)
    if nxd - nxh > 0.5:
        nxh += 1
    if nyd - nyh > 0.5:
        nyh += 1
    nx, ny = 2 * nxh + 1, 2 * nyh + 1
    xmax = ((2 * nxh + 1) * xstep - xstep) / 2
    ymax = ((2 * nyh + 1) * ystep - ystep) / 2
    psf = np.zeros((nx, ny))
    x = -xmax
    for i in range(nx):
        y = -ymax
        for j in range(ny):
            psf[i, j] = np.exp(-0.5 * ((x / sigma_x) ** 2 + (y / sigma_y) ** 2))
            y += ystep
        x += xstep
    return psf / psf.sum()


def moffat_psf(beta, fwhm, n, step):
    """Point-sampled circular Moffat (1 + r^2/alpha^2)^-beta on an n x n pixel grid, normalised."""
    alpha = fwhm / (2.0 * np.sqrt(2.0 ** (1.0 / beta) - 1.0))
    h = n // 2
    x = (np.arange(n) - h) * step
    r2 = x[:, None] ** 2 + x[None, :] ** 2
    psf = (1.0 + r2 / alpha ** 2) ** (-beta)
    return psf / psf.sum()


def c1():
    xv = corners(-2.5, 2.5, 100)
    return dict(
        name="C1", xvals=xv, yvals=xv.copy(), nsplit=2,
        lenses=[(SIE, (1.2, 0.8, 30.0, 0.0, 0.0, 0.0)), (SHEAR, (0.05, 20.0, 0.0, 0.0))],
        sources=[(1.0, 0.2, 1.5, 0.7, 0.0, 0.05, -0.03)],
        psf=gaussian_psf(0.1 * FWHM_TO_SIGMA, 0.02, 0.05, 0.05),
        sigma=0.01, seed=10,
        srcgrid=(-0.6, 0.6, -0.6, 0.6), src_n=(30, 30),
        regparam=1e3, integral_tol=1e-6,
    )


def c2_sources(seed=20251):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        r = 0.3 * np.sqrt(rng.uniform())
        ph = rng.uniform(0, 2 * np.pi)
        out.append((float(rng.uniform(0.5, 1.0)), float(rng.uniform(0.08, 0.2)), float(rng.uniform(0.8, 3.0)),
                    float(rng.uniform(0.6, 0.95)), float(rng.uniform(0, 180)),
                    float(r * np.cos(ph)), float(r * np.sin(ph))))
    return out


def c2(n_img=200, pix=0.04, n_src=60, src_half=1.0, psf_n=21):
    half = 0.5 * n_img * pix
    xv = corners(-half, half, n_img)
    return dict(
        name="C2", xvals=xv, yvals=xv.copy(), nsplit=2,
        lenses=[(NFW, (0.2, 5.0, 0.75, -15.0, 0.0, 0.0)),
                (SIE, (1.0, 0.85, 0.0, 0.0, 0.0, 0.0)),
                (SHEAR, (0.03, 0.0, 0.0, 0.0))],
        sources=c2_sources(),
        psf=moffat_psf(3.0, 0.12, psf_n, pix),
        sigma=0.01, seed=10,
        srcgrid=(-src_half, src_half, -src_half, src_half), src_n=(n_src, n_src),
        regparam=1e3, integral_tol=1e-6,
    )


def c2_small():
    """C2's physics at reduced size (oracle- and reference-friendly)."""
    return c2(n_img=60, pix=0.1333333333333333, n_src=24, src_half=1.0, psf_n=7)


def c3(n_img=400, nsplit=4):
    """Surrogate evaluation geometry: the numerical-profile lens of c2 (NFW + SIE + shear) over
    c1's 5"x5" field at 400x400 pixels with 4x4 supersampling (2.56 M rays), c1's source and
    30x30 source grid, c1's Gaussian PSF sampled at the finer pixel scale."""
    xv = corners(-2.5, 2.5, n_img)
    pix = 5.0 / n_img
    base = c1()
    return dict(
        name="C3", xvals=xv, yvals=xv.copy(), nsplit=nsplit,
        lenses=c2()["lenses"], sources=base["sources"],
        psf=gaussian_psf(0.1 * FWHM_TO_SIGMA, 0.02, pix, pix),
        sigma=0.01, seed=10, srcgrid=(-0.6, 0.6, -0.6, 0.6), src_n=(30, 30),
        regparam=1e3, integral_tol=1e-6,
    )


def c5(n_img=400, pix=0.02, n_src=100):
    """Mixed precision at scale: c2's lens and three-Sersic source over a 400x400 image at
    0.02"/pix (2x2 supersampling) on a 100x100 source grid (F is 10000 x 10000 dense), c2's
    Moffat PSF sampled at the finer pixel scale."""
    case = c2(n_img=n_img, pix=pix, n_src=n_src, src_half=1.0, psf_n=21)
    case["name"] = "C5"
    return case


def c_ns():
    """Orientation check: a non-square image (90 x 72 pixels of 0.05" x 0.045"), an anisotropic
    Gaussian PSF (sigma 0.06" in x, 0.035" in y, threshold 0.015, so 7 x 5 taps), an elliptical
    annulus mask with a rectangular hole, off-centre SIE + NFW + shear, and a non-square 28 x 22
    source grid.  No symmetry of the other workloads (square images, circular PSFs, full masks,
    square grids) can hide an x/y mix-up here."""
    xv = corners(-2.3, 2.2, 90)
    yv = corners(-1.62, 1.62, 72)
    xc = 0.5 * (xv[:-1] + xv[1:])
    yc = 0.5 * (yv[:-1] + yv[1:])
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    r = np.sqrt(X ** 2 + (Y / 0.8) ** 2)
    mask = (r > 0.45) & (r < 1.95) & ~((X > 0.9) & (X < 1.3) & (Y > -0.3) & (Y < 0.2))
    return dict(
        name="CNS", xvals=xv, yvals=yv, nsplit=2,
        lenses=[(SIE, (1.1, 0.7, 55.0, 0.05, -0.04, 0.0)), (NFW, (0.15, 3.0, 0.8, 20.0, 0.1, 0.02)),
                (SHEAR, (0.04, 75.0, 0.0, 0.0))],
        sources=[(1.0, 0.18, 1.3, 0.65, 40.0, 0.06, 0.04), (0.4, 0.1, 2.2, 0.8, -25.0, -0.2, 0.15)],
        psf=gaussian_psf_xy(0.06, 0.035, 0.015, 0.05, 0.045),
        psf_sigma=(0.06, 0.035), psf_threshold=0.015,
        mask=mask.astype(np.uint8),
        sigma=0.01, seed=10,
        srcgrid=(-0.75, 0.55, -0.45, 0.7), src_n=(28, 22),
        regparam=300.0, integral_tol=1e-6,
    )


def c_sple():
    """A power-law slope away from isothermal: SPLE alpha = 1.25 (the Tessore series) with the
    shear given as components, split_imgpixels off (one pixel-centre ray per pixel), a 64 x 56
    image of 0.05" pixels and a non-square 26 x 24 source grid."""
    xv = corners(-1.6, 1.6, 64)
    yv = corners(-1.4, 1.4, 56)
    return dict(
        name="CSPL", xvals=xv, yvals=yv, nsplit=0,
        lenses=[(SPLE, (1.05, 1.25, 0.75, 30.0, 0.05, -0.03)), (SHEAR_COMP, (0.04, -0.02, 0.0, 0.0))],
        sources=[(1.0, 0.15, 1.2, 0.7, 10.0, 0.04, 0.02)],
        psf=gaussian_psf(0.05, 0.01, 0.05, 0.05),
        sigma=0.01, seed=10,
        srcgrid=(-0.5, 0.5, -0.5, 0.5), src_n=(26, 24),
        regparam=100.0, integral_tol=1e-6,
    )


def c_sple_steep():
    """A steep SPLE (alpha = 1.6, q = 0.55: a long series) beside a spherical shallow one
    (alpha = 0.7), 3 x 3 subpixels, 48 x 48 image of 0.06" pixels."""
    xv = corners(-1.44, 1.44, 48)
    return dict(
        name="CSPS", xvals=xv, yvals=xv.copy(), nsplit=3,
        lenses=[(SPLE, (0.9, 1.6, 0.55, -30.0, 0.02, 0.04)), (SPLE, (0.25, 0.7, 1.0, 0.0, 0.6, -0.5)),
                (SHEAR, (0.03, 50.0, 0.0, 0.0))],
        sources=[(1.0, 0.12, 1.0, 0.8, -20.0, -0.03, 0.05)],
        psf=gaussian_psf(0.06, 0.02, 0.06, 0.06),
        sigma=0.01, seed=10,
        srcgrid=(-0.5, 0.5, -0.5, 0.5), src_n=(22, 22),
        regparam=50.0, integral_tol=1e-6,
    )


DEMO_FILES = ("sim_hst_image.fits", "mask.fits", "hst_psf.fits")
DEMO_TRUE = [(SPLE, (1.3634, 1.17163, 0.963867, 81.9, 0.0102892, 0.00358392)),
             (SHEAR_COMP, (0.0647257, -0.0575047, 0.0, 0.0))]
DEMO_START = [(SPLE, (1.35, 1.1, 0.85, 90.0, 0.0, 0.0)), (SHEAR_COMP, (0.05, 0.0, 0.0, 0.0))]


def demo():
    """The reference's shipped pixel-fit demo (delaunay_fit_demo.in / pixel_fit_demo.in) on a
    Cartesian source grid: the 81 x 81 HST-like image sim_hst_image.fits at data_pixel_size
    0.049", noise 0.00941285, the PSF hst_psf.fits cut at psf_threshold 1e-2, the mask mask.fits,
    split_imgpixels off, curvature regularization with optimize_regparam on over log10 lambda in
    [-5, 5], the source grid placed and sized automatically (auto_srcgrid, auto_src_npixels,
    pixel_fraction 0.3), and an SPLE (alpha 1.17) + shear-components lens.  The image, mask and
    PSF come from the files (`files`); DEMO_TRUE is the model that made the image, DEMO_START the
    demo's starting guess."""
    return dict(
        name="DEMO", files=DEMO_FILES, pixel_size=0.049, nsplit=0,
        lenses=DEMO_TRUE, sigma=0.00941285, psf_threshold=0.01,
        srcgrid=(-1.0, 1.0, -1.0, 1.0), src_n=(40, 40), auto_srcgrid=True, auto_src_npixels=True,
        optimize_regparam=True, regopt_bounds=(-5.0, 5.0),
        regparam=20.0, integral_tol=1e-6,
    )


def c1_geometry_200(src_side):
    """The 200 x 200 batched-evidence geometry (0.05"/pix, 10" field, 2 x 2 supersampling) with
    config 1's SIE + shear lens, Sersic source and 0.1" FWHM Gaussian PSF, on a src_side^2 grid."""
    base = c1()
    xv = corners(-5.0, 5.0, 200)
    return dict(
        name="C4s%d" % src_side, xvals=xv, yvals=xv.copy(), nsplit=2,
        lenses=base["lenses"], sources=base["sources"],
        psf=base["psf"], sigma=0.01, seed=10,
        srcgrid=(-0.6, 0.6, -0.6, 0.6), src_n=(src_side, src_side),
        regparam=1e3, integral_tol=1e-6,
    )


def proposals(case, n, seed=1234):
    """Seeded 1% Gaussian ball around the truth: SIE (b, q), NFW (ks, rs, q), shear gamma, SPLE
    (b, alpha, q), shear components."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m = []
        for t, p in case["lenses"]:
            p = list(p)
            idx = {SIE: (0, 1), NFW: (0, 1, 2), SHEAR: (0,), SPLE: (0, 1, 2), SHEAR_COMP: (0, 1)}[t]
            for i in idx:
                p[i] *= 1 + 0.01 * rng.standard_normal()
            if t in (SIE, NFW, SPLE):
                k = {SIE: 1, NFW: 2, SPLE: 2}[t]
                p[k] = min(p[k], 0.999)
            m.append((t, tuple(p)))
        out.append(m)
    return out


def c4(n=256, seed=4321):
    """Batched evidence: n proposals from a 1% Gaussian ball around config 1's truth (SIE + shear)
    on the 200 x 200 image of c1_geometry_200 (0.05"/pix, 2 x 2 supersampling), each with its own
    square source grid drawn uniformly from {20, 30, 40, 50, 60} cells per side (F up to
    3600 x 3600).  Returns (case, models, grids); the case's own grid is 60 x 60."""
    case = c1_geometry_200(60)
    case["name"] = "C4"
    rng = np.random.default_rng(seed)
    sides = rng.choice([20, 30, 40, 50, 60], size=n)
    grids = [(int(k), int(k)) for k in sides]
    return case, proposals(case, n, seed + 1), grids


def c4_nfw(n=256, seed=4321):
    """c4 on the numerical-profile workload instead: c2's image, PSF and NFW + SIE + shear truth."""
    case = c2()
    case["name"] = "C4nfw"
    rng = np.random.default_rng(seed)
    sides = rng.choice([20, 30, 40, 50, 60], size=n)
    grids = [(int(k), int(k)) for k in sides]
    return case, proposals(case, n, seed + 1), grids


WORKLOADS = {"C1": c1, "C2": c2, "C2s": c2_small, "C3": c3, "C5": c5, "CNS": c_ns}
