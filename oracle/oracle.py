"""ctypes wrapper for the CPU restatement (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, always as
the checker.  The product path (qlens_amd) never imports this module.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SIE, NFW, SHEAR, SPLE, SHEAR_COMP = 1, 2, 3, 4, 5


class OrcLens(C.Structure):
    _fields_ = [("type", C.c_int), ("p", C.c_double * 6)]


class OrcSersic(C.Structure):
    _fields_ = [("p", C.c_double * 7)]


class OrcProblem(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("xvals", C.POINTER(C.c_double)), ("yvals", C.POINTER(C.c_double)),
        ("data", C.POINTER(C.c_double)), ("covinv", C.POINTER(C.c_double)),
        ("mask", C.POINTER(C.c_ubyte)),
        ("nsplit", C.c_int),
        ("src_xmin", C.c_double), ("src_xmax", C.c_double),
        ("src_ymin", C.c_double), ("src_ymax", C.c_double),
        ("src_nx", C.c_int), ("src_ny", C.c_int),
        ("pnx", C.c_int), ("pny", C.c_int), ("psf", C.POINTER(C.c_double)),
        ("regparam", C.c_double),
        ("nlens", C.c_int), ("lenses", C.POINTER(OrcLens)),
        ("integral_tol", C.c_double),
    ]


class OrcResult(C.Structure):
    _fields_ = [
        ("minus2lnE", C.c_double), ("chisq0", C.c_double), ("logdetF", C.c_double),
        ("logdetR", C.c_double), ("Es", C.c_double), ("regterms", C.c_double),
        ("n_img", C.c_int), ("n_src", C.c_int), ("status", C.c_int),
        ("n_integrand_evals", C.c_longlong),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -f oracle/Makefile)")
        _LIB = C.CDLL(path)
        _LIB.orc_invert.restype = C.c_int
        _LIB.orc_mock_image.restype = C.c_int
        _LIB.orc_gaussian_psf.restype = C.c_int
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def lens_array(lenses):
    """lenses: list of (type, params) -> ctypes array"""
    arr = (OrcLens * max(1, len(lenses)))()
    for k, (t, p) in enumerate(lenses):
        arr[k].type = t
        for i, v in enumerate(p):
            arr[k].p[i] = float(v)
    return arr


def deflection(lenses, xy, tol=1e-3):
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
    out = np.zeros_like(xy)
    nev = C.c_longlong(0)
    la = lens_array(lenses)
    lib().orc_deflection(la, C.c_int(len(lenses)), C.c_double(tol), C.c_int(len(xy)), _dp(xy), _dp(out), C.byref(nev))
    return out, nev.value


def nfw_def_hess(lens_params, xy, tol=1e-6):
    """lens-frame (x, y) -> (ax, ay, psi_xx, psi_xy, psi_yy) for NFW params {ks, rs, q, ...}"""
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((len(xy), 5))
    la = lens_array([(NFW, lens_params)])
    lib().orc_nfw_def_hess(la, C.c_double(tol), C.c_int(len(xy)), _dp(xy), _dp(out))
    return out


def gaussian_psf(sigx, sigy, threshold, xstep, ystep):
    nx, ny = C.c_int(0), C.c_int(0)
    lib().orc_gaussian_psf(C.c_double(sigx), C.c_double(sigy), C.c_double(threshold), C.c_double(xstep),
                           C.c_double(ystep), C.byref(nx), C.byref(ny), None)
    psf = np.zeros((nx.value, ny.value))
    lib().orc_gaussian_psf(C.c_double(sigx), C.c_double(sigy), C.c_double(threshold), C.c_double(xstep),
                           C.c_double(ystep), C.byref(nx), C.byref(ny), _dp(psf))
    return psf


def normal_deviates(seed, n):
    out = np.zeros(n)
    lib().orc_normal_deviates(C.c_ulonglong(seed), C.c_int(n), _dp(out))
    return out


def mock_image(case):
    """Noisy mock image (x-major [nx, ny]) for a workload dict (see qlens_amd.workloads)."""
    xv, yv = case["xvals"], case["yvals"]
    nx, ny = len(xv) - 1, len(yv) - 1
    img = np.zeros((nx, ny))
    la = lens_array(case["lenses"])
    srcs = (OrcSersic * len(case["sources"]))()
    for k, p in enumerate(case["sources"]):
        for i, v in enumerate(p):
            srcs[k].p[i] = float(v)
    psf = np.ascontiguousarray(case["psf"], dtype=np.float64)
    lib().orc_mock_image(C.c_int(nx), C.c_int(ny), _dp(xv), _dp(yv), C.c_int(case["nsplit"]), la,
                         C.c_int(len(case["lenses"])), C.c_double(case["integral_tol"]), srcs,
                         C.c_int(len(case["sources"])), C.c_int(psf.shape[0]), C.c_int(psf.shape[1]),
                         _dp(psf), C.c_double(case["sigma"]), C.c_ulonglong(case["seed"]), _dp(img))
    return img


def invert(case, data, lenses=None, want_arrays=True):
    """Run the restated inversion; returns (result dict, arrays dict)."""
    xv = np.ascontiguousarray(case["xvals"], dtype=np.float64)
    yv = np.ascontiguousarray(case["yvals"], dtype=np.float64)
    nx, ny = len(xv) - 1, len(yv) - 1
    data = np.ascontiguousarray(data, dtype=np.float64)
    covinv = np.full((nx, ny), 1.0 / (case["sigma"] * case["sigma"]))
    if "covinv" in case:
        covinv = np.ascontiguousarray(case["covinv"], dtype=np.float64)
    psf = np.ascontiguousarray(case["psf"], dtype=np.float64)
    lenses = case["lenses"] if lenses is None else lenses
    la = lens_array(lenses)
    P = OrcProblem()
    P.nx, P.ny = nx, ny
    P.xvals, P.yvals = _dp(xv), _dp(yv)
    P.data, P.covinv = _dp(data), _dp(covinv)
    mask = case.get("mask")
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        P.mask = mask.ctypes.data_as(C.POINTER(C.c_ubyte))
    P.nsplit = case["nsplit"]
    P.src_xmin, P.src_xmax, P.src_ymin, P.src_ymax = case["srcgrid"]
    P.src_nx, P.src_ny = case["src_n"]
    P.pnx, P.pny = psf.shape
    P.psf = _dp(psf)
    P.regparam = case["regparam"]
    P.nlens = len(lenses)
    P.lenses = la
    P.integral_tol = case["integral_tol"]
    R = OrcResult()
    snx, sny = case["src_n"]
    ns = case["nsplit"]
    arrays = {}
    if want_arrays:
        arrays = dict(
            rays=np.zeros((nx, ny, max(ns * ns, 1), 2)), src=np.zeros((snx, sny)),
            active=np.zeros((snx, sny), dtype=np.int32), model=np.zeros((nx, ny)),
            mapped=np.zeros((nx, ny), dtype=np.uint8))
        lib().orc_invert(C.byref(P), C.byref(R), _dp(arrays["rays"]), _dp(arrays["src"]),
                         arrays["active"].ctypes.data_as(C.POINTER(C.c_int)), _dp(arrays["model"]),
                         arrays["mapped"].ctypes.data_as(C.POINTER(C.c_ubyte)))
    else:
        lib().orc_invert(C.byref(P), C.byref(R), None, None, None, None, None)
    res = {f: getattr(R, f) for f, _ in OrcResult._fields_}
    return res, arrays


def auto_source_grid(case, rays, data, auto_box=True, auto_npixels=True, pixel_fraction=0.3, noise_threshold=0.0):
    """Automatic source grid of one evaluation, from its subpixel rays (nx, ny, nsub, 2).

    Box (auto_box): ImagePixelGrid::find_optimal_sourcegrid (pixelgrid.cpp:10998-11080): over mask
    pixels whose data exceeds noise_threshold x noise (all of them when noise_threshold <= 0;
    window_size_for_sbavg is 0), the min / max of the mean subpixel source point (summed in
    subpixel order, divided by nsub), the limits at their defaults of +-1e30 (lens.cpp:381-384);
    then each side moved out by 5 % of the width.  Otherwise the case's srcgrid.
    Size (auto_npixels): find_optimal_sourcegrid_npixels (11469-11486): count = mask pixels whose
    centre ray (the pixel corner plus half the pixel size, 9659-9666, traced at the centre) lies
    strictly inside the box; nx0 = (int)sqrt(f count dx/dy), ny0 = (int)(nx0 dy/dx), both doubled
    (lens.cpp:12449-12450).  Otherwise the case's src_n.
    Returns (box [xmin, xmax, ymin, ymax], (snx, sny), count)."""
    xv = np.asarray(case["xvals"], dtype=np.float64)
    yv = np.asarray(case["yvals"], dtype=np.float64)
    nx, ny = len(xv) - 1, len(yv) - 1
    mask = case.get("mask")
    inmask = np.ones((nx, ny), dtype=bool) if mask is None else np.asarray(mask).astype(bool)
    if auto_box:
        use = inmask.copy()
        if noise_threshold > 0:
            sig = case["sigma"]
            use &= ~(np.asarray(data) <= noise_threshold * sig)
        nsub = rays.shape[2]
        xs = rays[:, :, 0, 0].copy()
        ys = rays[:, :, 0, 1].copy()
        for k in range(1, nsub):
            xs = xs + rays[:, :, k, 0]
            ys = ys + rays[:, :, k, 1]
        xs, ys = xs / nsub, ys / nsub
        xmin, xmax = float(xs[use].min()), float(xs[use].max())
        ymin, ymax = float(ys[use].min()), float(ys[use].max())
        xw, yw = 0.1 * (xmax - xmin), 0.1 * (ymax - ymin)
        box = [xmin - xw / 2, xmax + xw / 2, ymin - yw / 2, ymax + yw / 2]
    else:
        box = [float(v) for v in case["srcgrid"]]
    count = None
    if auto_npixels:
        pxl, pyl = (xv[-1] - xv[0]) / nx, (yv[-1] - yv[0]) / ny
        cx = xv[:-1] + 0.5 * pxl
        cy = yv[:-1] + 0.5 * pyl
        pts = np.stack(np.meshgrid(cx, cy, indexing="ij"), -1).reshape(-1, 2)
        alpha, _ = deflection(case["lenses"], pts, tol=case["integral_tol"])
        beta = (pts - alpha).reshape(nx, ny, 2)
        inside = (beta[..., 0] > box[0]) & (beta[..., 0] < box[1]) & (beta[..., 1] > box[2]) & (beta[..., 1] < box[3])
        count = int((inside & inmask).sum())
        dx, dy = box[1] - box[0], box[3] - box[2]
        nx0 = int(math.sqrt(pixel_fraction * count * dx / dy))
        ny0 = int(nx0 * dy / dx)
        size = (2 * nx0, 2 * ny0)
    else:
        size = tuple(case["src_n"])
    return box, size, count


# ---- on-disk formats (restated from the FITS standard and the reference's loaders) ----------
def read_fits_np(path):
    """Primary-HDU FITS image by the standard (NOST 100-2.0): 80-character cards in 2880-byte
    blocks up to END, then big-endian pixels with NAXIS1 fastest, value = BZERO + BSCALE raw (what
    cfitsio's fits_read_pix returns as doubles).  Returns (pix [naxis1, naxis2], header dict)."""
    raw = open(path, "rb").read()
    hdr, pos = {}, 0
    while True:
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8] == "=":
            hdr[key] = card[10:].split("/")[0].strip()
    data_off = -(-pos // 2880) * 2880
    bitpix, naxis = int(hdr["BITPIX"]), int(hdr["NAXIS"])
    n1 = int(hdr["NAXIS1"])
    n2 = int(hdr["NAXIS2"]) if naxis >= 2 else 1
    dt = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}[bitpix]
    a = np.frombuffer(raw, dtype=dt, count=n1 * n2, offset=data_off).astype(np.float64)
    bscale, bzero = float(hdr.get("BSCALE", 1.0)), float(hdr.get("BZERO", 0.0))
    if bscale != 1.0 or bzero != 0.0:
        a = a * bscale + bzero
    return a.reshape(n2, n1).T.copy(), hdr


def psf_crop(img, threshold):
    """QLens::load_psf_fits after the read (pixelgrid.cpp:6790-6840), restated: the index box of
    pixels above threshold x peak, widened to contain (nx/2, ny/2); half widths (box + 1)/2 (integer
    division); the window of 2 half + 1 pixels centred on (nx/2, ny/2); divided by its sum."""
    nx, ny = img.shape
    peak = img.max()
    ii, jj = np.nonzero(img > threshold * peak)
    imid, jmid = nx // 2, ny // 2
    imin, imax = min(imid, ii.min()), max(imid, ii.max())
    jmin, jmax = min(jmid, jj.min()), max(jmid, jj.max())
    hx, hy = (imax - imin + 1) // 2, (jmax - jmin + 1) // 2
    w = img[imid - hx:imid + hx + 1, jmid - hy:jmid + hy + 1].copy()
    norm = 0.0
    for i in range(w.shape[0]):
        for j in range(w.shape[1]):
            norm += w[i, j]
    return w / norm


def centred_grid(nx, ny, pixel_size):
    """load_data_fits pixel corners with a pixel size (pixelgrid.cpp:6198-6213): from -n s/2 in
    accumulated steps of s."""
    out = []
    for n in (nx, ny):
        v = -(0.5 * n * pixel_size)
        c = []
        for _ in range(n + 1):
            c.append(v)
            v += pixel_size
        out.append(np.array(c))
    return out[0], out[1]
