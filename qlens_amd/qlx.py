"""Python host mirror of QLens's pixel-inversion interface over the qlx C ABI (libqlx.so).

`PixelInversion` plays the role of the QLens object for this path: image data, PSF and
settings are loaded once, then `invert_image_surface_brightness_map(models)` returns -2 ln E for
each lens model, as QLens::invert_image_surface_brightness_map does (lens.cpp:12293-13195).
All numerical work runs in the HIP library; there is no CPU fallback: if libqlx.so is missing or
no GPU is visible, construction raises.
"""
import ctypes as C
import itertools
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqlx.so")

QLX_LENS_SIE, QLX_LENS_NFW, QLX_LENS_SHEAR, QLX_LENS_SPLE, QLX_LENS_SHEAR_COMP = 1, 2, 3, 4, 5
QLX_DEF_QUADRATURE, QLX_DEF_SURROGATE = 0, 1
QLX_PREC_FP64, QLX_PREC_MIXED = 0, 1
QLX_RAY_FP64, QLX_RAY_FP32 = 0, 1
QLX_MAX_LENSES = 8

STATUS = {0: "ok", 1: "some items failed", -1: "bad argument", -2: "bad state", -3: "HIP error",
          -4: "out of memory", -5: "not positive definite", -6: "no source pixels", -7: "out of range",
          -8: "refinement above tolerance"}
QLX_ITEMS_FAILED = 1
SENTINEL = 2e30  # minus2lnE of an item whose status is an error (lens.cpp:12453-12458)


class QlxLens(C.Structure):
    _fields_ = [("type", C.c_int), ("def_mode", C.c_int), ("p", C.c_double * 6)]


class QlxLensModel(C.Structure):
    _fields_ = [("nlens", C.c_int), ("lens", QlxLens * QLX_MAX_LENSES)]


class QlxOpts(C.Structure):
    _fields_ = [("device", C.c_int), ("hbm_budget", C.c_size_t)]


class QlxSettings(C.Structure):
    _fields_ = [("nsplit", C.c_int), ("src_xmin", C.c_double), ("src_xmax", C.c_double),
                ("src_ymin", C.c_double), ("src_ymax", C.c_double), ("src_nx", C.c_int), ("src_ny", C.c_int),
                ("regparam", C.c_double), ("integral_tol", C.c_double), ("precision", C.c_int),
                ("ray_precision", C.c_int), ("refine_iters", C.c_int), ("refine_tol", C.c_double),
                ("optimize_regparam", C.c_int), ("regopt_minlog", C.c_double), ("regopt_maxlog", C.c_double),
                ("regopt_tol", C.c_double), ("regopt_max_iter", C.c_int), ("auto_srcgrid", C.c_int),
                ("auto_src_npixels", C.c_int), ("pixel_fraction", C.c_double), ("noise_threshold", C.c_double),
                ("chol_panel", C.c_int), ("no_graphs", C.c_int)]


class QlxResult(C.Structure):
    _fields_ = [("minus2lnE", C.c_double), ("chisq0", C.c_double), ("logdetF", C.c_double),
                ("logdetR", C.c_double), ("Es", C.c_double), ("refine_resid", C.c_double),
                ("n_img_active", C.c_int), ("n_src_active", C.c_int), ("status", C.c_int),
                ("refine_steps", C.c_int), ("regparam", C.c_double), ("regopt_evals", C.c_int), ("src_nx", C.c_int),
                ("src_ny", C.c_int), ("pad_", C.c_int), ("src_xmin", C.c_double), ("src_xmax", C.c_double),
                ("src_ymin", C.c_double), ("src_ymax", C.c_double)]


class QlxSurrogateCfg(C.Structure):
    _fields_ = [("n_samples", C.c_int), ("steps", C.c_int), ("batch", C.c_int), ("lr", C.c_double),
                ("r_min", C.c_double), ("r_max", C.c_double), ("q_min", C.c_double), ("q_max", C.c_double),
                ("rs_min", C.c_double), ("rs_max", C.c_double), ("seed", C.c_ulonglong), ("tol", C.c_double),
                ("graph", C.c_int)]


class QlxFitParam(C.Structure):
    _fields_ = [("lens", C.c_int), ("param", C.c_int), ("lo", C.c_double), ("hi", C.c_double),
                ("transform", C.c_int), ("ta", C.c_double), ("tb", C.c_double), ("include_jacobian", C.c_int),
                ("prior", C.c_int), ("prior_mu", C.c_double), ("prior_sig", C.c_double), ("prior_norm", C.c_double)]


QLX_TRANSFORM_NONE, QLX_TRANSFORM_LOG10, QLX_TRANSFORM_LINEAR = 0, 1, 2
QLX_PRIOR_UNIFORM, QLX_PRIOR_LOG, QLX_PRIOR_GAUSS = 0, 1, 2

_LIB = None


def load_library(path=None):
    """Load libqlx.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("libqlx.so not found at %s; build it with `python -m qlens_amd.build`" % p)
    lib = C.CDLL(p)
    lib.qlx_last_error.restype = C.c_char_p
    lib.qlx_last_error.argtypes = [C.c_void_p]
    for name in ["qlx_ctx_create", "qlx_ctx_destroy", "qlx_set_image", "qlx_set_psf", "qlx_set_psf_gaussian",
                 "qlx_set_settings", "qlx_eval", "qlx_eval_batch_grids", "qlx_get_source", "qlx_raytrace",
                 "qlx_nfw_def_hess", "qlx_mock_image", "qlx_kernel_time", "qlx_selftest_mfma",
                 "qlx_surrogate_train", "qlx_surrogate_eval", "qlx_surrogate_grad", "qlx_surrogate_param_count",
                 "qlx_surrogate_get_params", "qlx_surrogate_set_params", "qlx_surrogate_save", "qlx_surrogate_load",
                 "qlx_get_stats", "qlx_set_fit_model", "qlx_loglike"]:
        getattr(lib, name).restype = C.c_int
    lib.qlx_get_rays.restype = C.c_int
    lib.qlx_get_rays.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_size_t]
    lib.qlx_get_psf.restype = C.c_int
    lib.qlx_get_psf.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_size_t]
    lib.qlx_get_source.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                   C.POINTER(C.c_int), C.c_size_t, C.POINTER(C.c_double), C.c_size_t]
    lib.qlx_io_last_error.restype = C.c_char_p
    lib.qlx_io_last_error.argtypes = []
    pint, pdbl = C.POINTER(C.c_int), C.POINTER(C.c_double)
    for name, args in (("qlx_read_image_text", [C.c_char_p, pint, pint, pdbl, pdbl, pdbl, C.c_size_t]),
                       ("qlx_read_fits", [C.c_char_p, pint, pint, pdbl, C.c_size_t, pdbl, pdbl]),
                       ("qlx_pixel_grid", [C.c_int, C.c_int, C.c_double, pdbl, pdbl]),
                       ("qlx_psf_from_image", [pdbl, C.c_int, C.c_int, C.c_double, pint, pint, pdbl, C.c_size_t]),
                       ("qlx_mask_from_image", [pdbl, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_ubyte)])):
        getattr(lib, name).restype = C.c_int
        getattr(lib, name).argtypes = args
    if path is None:
        _LIB = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# numpy views of the ctypes arrays a batch call fills and reads (explicit layouts, checked against
# ctypes' own): a 256-model batch is packed and unpacked in a fraction of the per-field time
_LENS_DT = np.dtype([("type", "<i4"), ("def_mode", "<i4"), ("p", "<f8", (6,))])
_MODEL_DT = np.dtype({"names": ["nlens", "lens"], "formats": ["<i4", (_LENS_DT, (QLX_MAX_LENSES,))],
                      "offsets": [0, QlxLensModel.lens.offset], "itemsize": C.sizeof(QlxLensModel)})
_RESULT_DT = np.dtype({"names": [f for f, _ in QlxResult._fields_],
                       "formats": ["<f8" if t is C.c_double else "<i4" for _, t in QlxResult._fields_],
                       "offsets": [getattr(QlxResult, f).offset for f, _ in QlxResult._fields_],
                       "itemsize": C.sizeof(QlxResult)})
assert _LENS_DT.itemsize == C.sizeof(QlxLens)


def _fill_same_layout(a, models, def_mode):
    """make_models' common case: every model has the first one's lens types and parameter counts
    (a chain's proposals), so each lens position is filled from one array.  Returns False, having
    written nothing, when the models differ in layout."""
    n, nl = len(models), len(models[0])
    if nl > QLX_MAX_LENSES or set(map(len, models)) != {nl}:
        return False
    cols = []
    for k in range(nl):
        col = [lm[k] for lm in models]
        t0 = col[0][0]
        if [c[0] for c in col].count(t0) != n:
            return False
        prm = [c[1] for c in col]
        m = len(prm[0])
        if m > a["lens"]["p"].shape[-1] or set(map(len, prm)) != {m}:
            return False
        cols.append((t0, np.fromiter(itertools.chain.from_iterable(prm), dtype=np.float64, count=n * m).reshape(n, m)))
    a["nlens"][:n] = nl
    L = a["lens"]
    for k, (t, p) in enumerate(cols):
        L["type"][:n, k] = t
        L["def_mode"][:n, k] = def_mode if t == QLX_LENS_NFW else QLX_DEF_QUADRATURE
        L["p"][:n, k, :p.shape[1]] = p
    return True


def make_models(models, def_mode=QLX_DEF_QUADRATURE):
    """[[(type, params)], ...] in QLens conventions -> ctypes array of QlxLensModel"""
    n = len(models)
    arr = (QlxLensModel * max(1, n))()
    a = np.frombuffer(arr, dtype=_MODEL_DT)
    if n and _fill_same_layout(a, models, def_mode):
        return arr
    a["nlens"][:n] = [len(lm) for lm in models]
    flat = [(i, k, t, prm) for i, lm in enumerate(models) for k, (t, prm) in enumerate(lm)]
    if not flat:
        return arr
    mi = np.fromiter((f[0] for f in flat), dtype=np.intp, count=len(flat))
    li = np.fromiter((f[1] for f in flat), dtype=np.intp, count=len(flat))
    if li.max() >= QLX_MAX_LENSES:
        raise IndexError("at most %d lenses per model" % QLX_MAX_LENSES)
    ty = np.fromiter((f[2] for f in flat), dtype=np.int32, count=len(flat))
    L = a["lens"]
    L["type"][mi, li] = ty
    L["def_mode"][mi, li] = np.where(ty == QLX_LENS_NFW, def_mode, QLX_DEF_QUADRATURE)
    bylen = {}
    for j, f in enumerate(flat):
        bylen.setdefault(len(f[3]), []).append(j)
    for m, js in bylen.items():
        sel = np.asarray(js, dtype=np.intp)
        L["p"][mi[sel], li[sel], :m] = np.asarray([flat[j][3] for j in js], dtype=np.float64).reshape(-1, m)
    return arr


def make_model(lenses, def_mode=QLX_DEF_QUADRATURE):
    """lenses: [(type, params)] in QLens conventions -> QlxLensModel"""
    m = QlxLensModel()
    m.nlens = len(lenses)
    for k, (t, p) in enumerate(lenses):
        m.lens[k].type = t
        m.lens[k].def_mode = def_mode if t == QLX_LENS_NFW else QLX_DEF_QUADRATURE
        for i, v in enumerate(p):
            m.lens[k].p[i] = float(v)
    return m


class QlxError(RuntimeError):
    pass


# -- on-disk inputs (host only: no context or device) ------------------------------------------
def _io(rc, what):
    if rc != 0:
        raise QlxError("%s failed (%s): %s" % (what, STATUS.get(rc, rc), load_library().qlx_io_last_error().decode()))


def read_image_text(root):
    """ImagePixelData::load_data (pixelgrid.cpp:5797-5904): root.x / root.y corners and root.dat
    surface brightness -> (xvals, yvals, sb [nx, ny])."""
    lib = load_library()
    nx, ny = C.c_int(0), C.c_int(0)
    _io(lib.qlx_read_image_text(root.encode(), C.byref(nx), C.byref(ny), None, None, None, 0), "qlx_read_image_text")
    xv, yv, sb = np.zeros(nx.value + 1), np.zeros(ny.value + 1), np.zeros((nx.value, ny.value))
    _io(lib.qlx_read_image_text(root.encode(), C.byref(nx), C.byref(ny), _dp(xv), _dp(yv), _dp(sb), sb.size),
        "qlx_read_image_text")
    return xv, yv, sb


def read_fits(path):
    """Primary FITS image as the reference reads it through cfitsio -> (pix [naxis1, naxis2],
    PXSIZE or None, PXNOISE or None)."""
    lib = load_library()
    n1, n2 = C.c_int(0), C.c_int(0)
    ps, pn = C.c_double(0), C.c_double(0)
    _io(lib.qlx_read_fits(path.encode(), C.byref(n1), C.byref(n2), None, 0, C.byref(ps), C.byref(pn)), "qlx_read_fits")
    pix = np.zeros((n1.value, n2.value))
    _io(lib.qlx_read_fits(path.encode(), C.byref(n1), C.byref(n2), _dp(pix), pix.size, None, None), "qlx_read_fits")
    return pix, (ps.value if np.isfinite(ps.value) else None), (pn.value if np.isfinite(pn.value) else None)


def pixel_grid(nx, ny, pixel_size):
    """Corners of an nx x ny image of square pixels centred on the origin (load_data_fits with a
    pixel size, pixelgrid.cpp:6198-6213)."""
    xv, yv = np.zeros(nx + 1), np.zeros(ny + 1)
    _io(load_library().qlx_pixel_grid(int(nx), int(ny), C.c_double(pixel_size), _dp(xv), _dp(yv)), "qlx_pixel_grid")
    return xv, yv


def psf_from_image(img, threshold=0.01):
    """The PSF matrix load_psf_fits keeps from an input image (pixelgrid.cpp:6790-6840)."""
    lib = load_library()
    img = np.ascontiguousarray(img, dtype=np.float64)
    pnx, pny = C.c_int(0), C.c_int(0)
    _io(lib.qlx_psf_from_image(_dp(img), img.shape[0], img.shape[1], C.c_double(threshold), C.byref(pnx), C.byref(pny),
                               None, 0), "qlx_psf_from_image")
    psf = np.zeros((pnx.value, pny.value))
    _io(lib.qlx_psf_from_image(_dp(img), img.shape[0], img.shape[1], C.c_double(threshold), C.byref(pnx), C.byref(pny),
                               _dp(psf), psf.size), "qlx_psf_from_image")
    return psf


def mask_from_image(m, nx, ny, mask=None):
    """Apply a mask image (nonzero = in) to an nx x ny mask as load_mask_fits does
    (pixelgrid.cpp:6532-6620); mask defaults to every pixel in."""
    m = np.ascontiguousarray(m, dtype=np.float64)
    out = np.ones((nx, ny), dtype=np.uint8) if mask is None else np.array(mask, dtype=np.uint8, copy=True).reshape(nx, ny)
    _io(load_library().qlx_mask_from_image(_dp(m), m.shape[0], m.shape[1], int(nx), int(ny),
                                           out.ctypes.data_as(C.POINTER(C.c_ubyte))), "qlx_mask_from_image")
    return out


class PixelInversion:
    """One qlx context (one device stream) holding the image, PSF and inversion settings."""

    def __init__(self, device=0, hbm_budget=0):
        """hbm_budget: bytes the context may use for batch admission (0 = 85% of free HBM); a
        smaller budget splits a call into more waves (same results)."""
        self.lib = load_library()
        self.ctx = C.c_void_p()
        opts = QlxOpts(device, int(hbm_budget))
        rc = self.lib.qlx_ctx_create(C.byref(opts), C.byref(self.ctx))
        if rc != 0:
            raise QlxError("qlx_ctx_create failed (%s): %s" % (STATUS.get(rc, rc), self.lib.qlx_last_error(None).decode()))
        self.nx = self.ny = 0
        self.settings = None

    def close(self):
        if self.ctx:
            self.lib.qlx_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise QlxError("%s failed (%s): %s" % (what, STATUS.get(rc, rc), self.lib.qlx_last_error(self.ctx).decode()))

    # -- inputs ------------------------------------------------------------------------------
    def load_image(self, xvals, yvals, data, sigma=None, covinv=None, mask=None):
        """ImagePixelData equivalent: corners + x-major data; noise as sigma or a covinv map."""
        xv = np.ascontiguousarray(xvals, dtype=np.float64)
        yv = np.ascontiguousarray(yvals, dtype=np.float64)
        nx, ny = len(xv) - 1, len(yv) - 1
        d = np.ascontiguousarray(data, dtype=np.float64).reshape(nx, ny)
        if covinv is None:
            covinv = np.full((nx, ny), 1.0 / (sigma * sigma))
        ci = np.ascontiguousarray(covinv, dtype=np.float64).reshape(nx, ny)
        mk = None
        if mask is not None:
            mk = np.ascontiguousarray(mask, dtype=np.uint8).reshape(nx, ny)
        self._check(self.lib.qlx_set_image(self.ctx, nx, ny, _dp(xv), _dp(yv), _dp(d), _dp(ci),
                                           mk.ctypes.data_as(C.POINTER(C.c_ubyte)) if mk is not None else None),
                    "qlx_set_image")
        self.nx, self.ny = nx, ny
        self._keep = (xv, yv, d, ci, mk)

    def load_image_text(self, root, sigma):
        """'sbmap loadimg <root>' on a text image with uniform pixel noise sigma."""
        xv, yv, sb = read_image_text(root)
        self.load_image(xv, yv, sb, sigma=sigma)

    def load_image_fits(self, path, pixel_size=None, bounds=None, sigma=None):
        """'sbmap loadimg <file.fits>' (load_data_fits, pixelgrid.cpp:6016-6280): the pixel size
        from the file's PXSIZE keyword, else `pixel_size` (data_pixel_size), gives a grid centred
        on the origin; otherwise `bounds` (xmin, xmax, ymin, ymax) are split in accumulated steps.
        The noise is `sigma`, else the file's PXNOISE.  The mask is every pixel."""
        pix, pxs, pxn = read_fits(path)
        nx, ny = pix.shape
        ps = pxs if pxs is not None else pixel_size
        if ps is not None:
            xv, yv = pixel_grid(nx, ny, ps)
        elif bounds is not None:
            xv, yv = np.zeros(nx + 1), np.zeros(ny + 1)
            for arr, lo, hi, n in ((xv, bounds[0], bounds[1], nx), (yv, bounds[2], bounds[3], ny)):
                step, v = (hi - lo) / n, lo
                for i in range(n + 1):
                    arr[i] = v
                    v += step
        else:
            raise QlxError("load_image_fits: no PXSIZE in %s; give pixel_size or bounds" % path)
        sig = sigma if sigma is not None else pxn
        if sig is None:
            raise QlxError("load_image_fits: no PXNOISE in %s; give sigma" % path)
        self.load_image(xv, yv, pix, sigma=sig)

    def load_mask_fits(self, path):
        """'sbmap loadmask <file.fits>' into the primary mask (load_mask_fits)."""
        m, _, _ = read_fits(path)
        xv, yv, d, ci, mk = self._keep
        mask = mask_from_image(m, self.nx, self.ny, mk)
        self.load_image(xv, yv, d, covinv=ci, mask=mask)

    def load_psf_fits(self, path, threshold=0.01):
        """'sbmap loadpsf <file.fits>' with psf_threshold (load_psf_fits)."""
        m, _, _ = read_fits(path)
        self.set_psf(psf_from_image(m, threshold))

    def set_psf(self, psf):
        psf = np.ascontiguousarray(psf, dtype=np.float64)
        self._check(self.lib.qlx_set_psf(self.ctx, psf.shape[0], psf.shape[1], _dp(psf)), "qlx_set_psf")

    def set_psf_gaussian(self, sigma_x, sigma_y, threshold):
        self._check(self.lib.qlx_set_psf_gaussian(self.ctx, C.c_double(sigma_x), C.c_double(sigma_y),
                                                  C.c_double(threshold)), "qlx_set_psf_gaussian")

    def set_settings(self, nsplit, srcgrid, src_n, regparam, integral_tol=1e-3, precision=QLX_PREC_FP64,
                     ray_precision=QLX_RAY_FP64, refine_iters=0, refine_tol=0.0, optimize_regparam=False,
                     regopt_bounds=(-3.0, 5.0), regopt_tol=0.01, regopt_max_iter=20, auto_srcgrid=False,
                     auto_src_npixels=False, pixel_fraction=0.3, noise_threshold=0.0, chol_panel=0, graphs=True):
        """QLens settings of the inversion; optimize_regparam runs Brent over log10(lambda) in
        regopt_bounds per item (the reference's 'optimize_regparam on', FP64 only); auto_srcgrid and
        auto_src_npixels find each item's source-grid box and size from its own rays (the
        reference's 'auto_srcgrid on' / 'auto_src_npixels on'; srcgrid / src_n are then unused)."""
        s = QlxSettings(nsplit, srcgrid[0], srcgrid[1], srcgrid[2], srcgrid[3], src_n[0], src_n[1], regparam,
                        integral_tol, precision, ray_precision, refine_iters, refine_tol,
                        1 if optimize_regparam else 0, regopt_bounds[0], regopt_bounds[1], regopt_tol,
                        int(regopt_max_iter), 1 if auto_srcgrid else 0, 1 if auto_src_npixels else 0,
                        pixel_fraction, noise_threshold, int(chol_panel), 0 if graphs else 1)
        self._check(self.lib.qlx_set_settings(self.ctx, C.byref(s)), "qlx_set_settings")
        self.settings = dict(nsplit=nsplit, srcgrid=srcgrid, src_n=src_n, regparam=regparam,
                             integral_tol=integral_tol, precision=precision)

    def load_workload(self, case, data, precision=QLX_PREC_FP64, ray_precision=QLX_RAY_FP64, optimize_regparam=False):
        self.load_image(case["xvals"], case["yvals"], data, sigma=case["sigma"], mask=case.get("mask"))
        self.set_psf(case["psf"])
        self.set_settings(case["nsplit"], case["srcgrid"], case["src_n"], case["regparam"], case["integral_tol"],
                          precision=precision, ray_precision=ray_precision,
                          optimize_regparam=optimize_regparam or bool(case.get("optimize_regparam")),
                          regopt_bounds=case.get("regopt_bounds", (-3.0, 5.0)),
                          auto_srcgrid=bool(case.get("auto_srcgrid")), auto_src_npixels=bool(case.get("auto_src_npixels")))

    # -- evaluation --------------------------------------------------------------------------
    def invert_image_surface_brightness_map(self, lens_models, def_mode=QLX_DEF_QUADRATURE, src_n=None):
        """Evaluate -2 ln E for one lens model ([(type, params)]) or a list of them.
        Returns a list of result dicts (one per model)."""
        single = len(lens_models) > 0 and isinstance(lens_models[0], tuple) and isinstance(lens_models[0][0], int)
        models = [lens_models] if single else list(lens_models)
        arr = make_models(models, def_mode)
        res = (QlxResult * max(1, len(models)))()
        if src_n is None:
            rc = self.lib.qlx_eval(self.ctx, len(models), arr, res)
        else:
            sn = np.ascontiguousarray(np.asarray(src_n, dtype=np.int32).reshape(-1, 2))
            rc = self.lib.qlx_eval_batch_grids(self.ctx, len(models), arr, sn.ctypes.data_as(C.POINTER(C.c_int)), res)
        if rc < 0:
            self._check(rc, "qlx_eval")
        self._last_n = len(models)
        names = _RESULT_DT.names
        return [dict(zip(names, row)) for row in np.frombuffer(res, dtype=_RESULT_DT)[:len(models)].tolist()]

    def get_source(self, item=-1):
        """Item `item` of the last evaluation (negative: from the end), as update_source_amplitudes
        leaves it (pixelgrid.cpp:17850-17881): dict with 'src' [snx, sny] (0 off the active set),
        'active_index' [snx, sny] (-1 inactive) and 'model' [nx, ny]."""
        if item < 0:
            item += getattr(self, "_last_n", 0)
        snx, sny = C.c_int(0), C.c_int(0)
        self._check(self.lib.qlx_get_source(self.ctx, int(item), C.byref(snx), C.byref(sny), None, None, 0, None, 0),
                    "qlx_get_source")
        src = np.zeros((snx.value, sny.value))
        act = np.zeros((snx.value, sny.value), dtype=np.int32)
        model = np.zeros((self.nx, self.ny))
        self._check(self.lib.qlx_get_source(self.ctx, int(item), None, None, _dp(src),
                                            act.ctypes.data_as(C.POINTER(C.c_int)), src.size, _dp(model), model.size),
                    "qlx_get_source")
        return {"src": src, "active_index": act, "model": model}

    # -- fit model and likelihood (QLens::fitmodel_loglike_extended_source) -----------------------
    def set_fit_model(self, base_lenses, params):
        """base_lenses: [(type, params)]; params: one dict per varied parameter with keys lens, param
        and optionally lo, hi (penalty limits on the physical value), transform ('none', 'log10',
        ('linear', a, b)), jacobian (bool), prior ('uniform', ('log', norm), ('gauss', mu, sig, norm))."""
        arr = (QlxFitParam * max(1, len(params)))()
        for k, d in enumerate(params):
            f = arr[k]
            f.lens, f.param = int(d["lens"]), int(d["param"])
            f.lo, f.hi = float(d.get("lo", 0.0)), float(d.get("hi", 0.0))
            t = d.get("transform", "none")
            if t == "log10":
                f.transform = QLX_TRANSFORM_LOG10
            elif isinstance(t, tuple) and t[0] == "linear":
                f.transform, f.ta, f.tb = QLX_TRANSFORM_LINEAR, float(t[1]), float(t[2])
            f.include_jacobian = 1 if d.get("jacobian") else 0
            pr = d.get("prior", "uniform")
            if isinstance(pr, tuple) and pr[0] == "log":
                f.prior, f.prior_norm = QLX_PRIOR_LOG, float(pr[1])
            elif isinstance(pr, tuple) and pr[0] == "gauss":
                f.prior, f.prior_mu, f.prior_sig, f.prior_norm = QLX_PRIOR_GAUSS, float(pr[1]), float(pr[2]), float(pr[3])
        m = make_model(base_lenses)
        self._check(self.lib.qlx_set_fit_model(self.ctx, C.byref(m), len(params), arr), "qlx_set_fit_model")
        self._nfit = len(params)

    def loglike(self, x):
        """-ln L for sampler points x[n, nparam] in one batched call; returns (loglike[n], results)."""
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, self._nfit))
        n = len(x)
        ll = np.zeros(n)
        res = (QlxResult * max(1, n))()
        rc = self.lib.qlx_loglike(self.ctx, n, _dp(x), _dp(ll), res)
        if rc < 0:
            self._check(rc, "qlx_loglike")
        self._last_n = n
        names = _RESULT_DT.names
        return ll, [dict(zip(names, row)) for row in np.frombuffer(res, dtype=_RESULT_DT)[:n].tolist()]

    def get_rays(self, item=-1):
        """Subpixel source-plane points of an item of the last evaluation: [nx, ny, nsplit^2, 2]
        (one pixel-centre ray per pixel when nsplit is 0)."""
        if item < 0:
            item += getattr(self, "_last_n", 0)
        ns = self.settings["nsplit"]
        out = np.zeros((self.nx, self.ny, max(ns * ns, 1), 2))
        self._check(self.lib.qlx_get_rays(self.ctx, int(item), _dp(out), out.size), "qlx_get_rays")
        return out

    def get_psf(self):
        """The PSF matrix in use, [pnx, pny]."""
        nx, ny = C.c_int(0), C.c_int(0)
        self._check(self.lib.qlx_get_psf(self.ctx, C.byref(nx), C.byref(ny), None, 0), "qlx_get_psf")
        psf = np.zeros((nx.value, ny.value))
        self._check(self.lib.qlx_get_psf(self.ctx, None, None, _dp(psf), psf.size), "qlx_get_psf")
        return psf

    def get_solution(self, item=-1):
        """(source amplitudes [snx, sny], model image [nx, ny]) of an item of the last evaluation."""
        r = self.get_source(item)
        return r["src"], r["model"]

    def raytrace(self, lenses, xy):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        out = np.zeros_like(xy)
        m = make_model(lenses)
        self._check(self.lib.qlx_raytrace(self.ctx, C.byref(m), len(xy), _dp(xy), _dp(out)), "qlx_raytrace")
        return out

    def nfw_def_hess(self, nfw_params, xy, tol=1e-6):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        out = np.zeros((len(xy), 5))
        p = np.ascontiguousarray(nfw_params, dtype=np.float64)
        self._check(self.lib.qlx_nfw_def_hess(self.ctx, _dp(p), C.c_double(tol), len(xy), _dp(xy), _dp(out)),
                    "qlx_nfw_def_hess")
        return out

    def mock_image(self, lenses, sersics, sigma, seed):
        out = np.zeros((self.nx, self.ny))
        s = np.ascontiguousarray(np.asarray(sersics, dtype=np.float64).reshape(-1, 7))
        m = make_model(lenses)
        self._check(self.lib.qlx_mock_image(self.ctx, C.byref(m), len(s), _dp(s), C.c_double(sigma),
                                            C.c_ulonglong(seed), _dp(out)), "qlx_mock_image")
        return out

    # -- learned deflection surrogate -----------------------------------------------------------
    def surrogate_train(self, n_samples=2000000, steps=4000, batch=8192, lr=2e-3, r_box=(1e-3, 20.0), q_box=(0.3, 1.0),
                        rs_box=(1.0, 20.0), seed=7, tol=1e-8, graph=True):
        """Generate quadrature targets on the GPU and train the network there; returns the final
        minibatch loss (standardized outputs)."""
        cfg = QlxSurrogateCfg(int(n_samples), int(steps), int(batch), float(lr), r_box[0], r_box[1], q_box[0], q_box[1],
                              rs_box[0], rs_box[1], int(seed), float(tol), 1 if graph else 0)
        loss = C.c_double(0)
        self._check(self.lib.qlx_surrogate_train(self.ctx, C.byref(cfg), C.byref(loss)), "qlx_surrogate_train")
        return loss.value

    def surrogate_eval(self, x4):
        """(ln r, phi, q, ln rs)[n, 4] in the NFW lens frame -> (alpha_x, alpha_y, psi_xx, psi_xy, psi_yy) / ks."""
        x = np.ascontiguousarray(x4, dtype=np.float32).reshape(-1, 4)
        y = np.zeros((len(x), 5), dtype=np.float32)
        self._check(self.lib.qlx_surrogate_eval(self.ctx, len(x), x.ctypes.data_as(C.POINTER(C.c_float)),
                                                y.ctypes.data_as(C.POINTER(C.c_float))), "qlx_surrogate_eval")
        return y

    def surrogate_grad(self, X, T):
        """Loss and gradient of the network on standardized features X[n, 4] and targets T[n, 5]."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 4)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, 5)
        g = np.zeros(self.lib.qlx_surrogate_param_count(self.ctx), dtype=np.float32)
        loss = C.c_float(0)
        self._check(self.lib.qlx_surrogate_grad(self.ctx, len(X), X.ctypes.data_as(C.POINTER(C.c_float)),
                                                T.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss),
                                                g.ctypes.data_as(C.POINTER(C.c_float))), "qlx_surrogate_grad")
        return loss.value, g

    def surrogate_params(self):
        p = np.zeros(self.lib.qlx_surrogate_param_count(self.ctx), dtype=np.float32)
        self._check(self.lib.qlx_surrogate_get_params(self.ctx, p.ctypes.data_as(C.POINTER(C.c_float))), "qlx_surrogate_get_params")
        return p

    def set_surrogate_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.lib.qlx_surrogate_set_params(self.ctx, p.ctypes.data_as(C.POINTER(C.c_float))), "qlx_surrogate_set_params")

    def surrogate_save(self, path):
        self._check(self.lib.qlx_surrogate_save(self.ctx, str(path).encode()), "qlx_surrogate_save")

    def surrogate_load(self, path):
        self._check(self.lib.qlx_surrogate_load(self.ctx, str(path).encode()), "qlx_surrogate_load")

    def kernel_time(self, name):
        ms = C.c_double(0)
        n = C.c_int(0)
        rc = self.lib.qlx_kernel_time(self.ctx, name.encode(), C.byref(ms), C.byref(n))
        return (ms.value, n.value) if rc == 0 else (None, 0)

    def selftest_mfma(self, kind=0):
        """kind 0: f64 16x16x4, 1: f32 16x16x4, 2: bf16 16x16x32, 3: f16 16x16x32 (see include/qlx.h)."""
        out = np.zeros(256)
        self._check(self.lib.qlx_selftest_mfma(self.ctx, int(kind), _dp(out)), "qlx_selftest_mfma")
        return out


class QlxStats(C.Structure):
    _fields_ = [("syrk_dense_flops", C.c_double), ("syrk_exec_flops", C.c_double), ("syrk_alg_flops", C.c_double),
                ("lb_bytes", C.c_double), ("integrand_evals", C.c_double), ("chol_flops", C.c_double),
                ("n_items", C.c_int), ("n_img", C.c_int), ("n_src", C.c_int), ("ld_src", C.c_int),
                ("rows_pad", C.c_int), ("n_waves", C.c_int), ("chol_panel", C.c_int),
                ("graph_captures", C.c_int), ("graph_replays", C.c_int)]


def get_stats(inv):
    """Work counters of the last wave of the last evaluation (qlx_get_stats)."""
    s = QlxStats()
    inv._check(inv.lib.qlx_get_stats(inv.ctx, C.byref(s)), "qlx_get_stats")
    return {f: getattr(s, f) for f, _ in QlxStats._fields_}
