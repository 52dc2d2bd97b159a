"""On-disk inputs (host only, no device): the text image format, FITS images (data, mask, PSF)
and the transformations the reference applies after reading them, against the oracle's
restatements (oracle.read_fits_np from the FITS standard, psf_crop, centred_grid) and against
data whose values are known by construction.

The FITS reader of the reference is cfitsio, which this image does not have; the fixtures are
the FITS files the reference ships for its demo (tests/golden/fits/: sim_hst_image.fits,
mask.fits, hst_psf.fits).  The orientation (axis 1 = x) is pinned end to end on the GPU by the
demo's chi-square at the true lens (test_gpu_parity.py)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from qlens_amd import qlx
from qlens_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
FITS = os.path.join(HERE, "golden", "fits")
LIB = os.path.join(os.path.dirname(HERE), "qlens_amd", "libqlx.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libqlx.so not built")


def _write_text_image(root, xv, yv, data):
    # the layout the reference reads (one value per line for corners, rows of ny values for the
    # data, 17 significant digits: the same doubles come back)
    with open(root + ".x", "w") as f:
        f.write("\n".join("%.17g" % v for v in xv) + "\n")
    with open(root + ".y", "w") as f:
        f.write("  ".join("%.17g" % v for v in yv) + "\n")  # any whitespace separates values
    with open(root + ".dat", "w") as f:
        for row in data:
            f.write(" ".join("%.17g" % v for v in row) + "\n")


def test_text_image_round_trip(tmp_path, golden_dir):
    g = np.load(os.path.join(golden_dir, "inversion_CNS.npz"))
    case = W.c_ns()
    root = str(tmp_path / "img")
    _write_text_image(root, case["xvals"], case["yvals"], g["data"])
    xv, yv, sb = qlx.read_image_text(root)
    assert np.array_equal(xv, case["xvals"]) and np.array_equal(yv, case["yvals"])
    assert sb.shape == (90, 72) and np.array_equal(sb, g["data"])


def test_text_image_errors(tmp_path):
    root = str(tmp_path / "bad")
    _write_text_image(root, [0, 1, 2], [0, 1], np.ones((2, 1)))
    with open(root + ".dat", "w") as f:
        f.write("1.0\n")  # one value for a 2 x 1 image
    with pytest.raises(qlx.QlxError):
        qlx.read_image_text(root)
    with pytest.raises(qlx.QlxError):
        qlx.read_image_text(str(tmp_path / "missing"))


@pytest.mark.parametrize("name,shape", [("sim_hst_image.fits", (81, 81)), ("mask.fits", (81, 81)),
                                        ("hst_psf.fits", (37, 37))])
def test_shipped_fits_read_as_the_standard_says(name, shape):
    path = os.path.join(FITS, name)
    pix, pxs, pxn = qlx.read_fits(path)
    ref, hdr = O.read_fits_np(path)
    assert pix.shape == shape and np.array_equal(pix, ref)
    assert pxs is None and pxn is None


def _write_fits(path, arr_xy, bitpix, extra_cards=(), bscale=None, bzero=None):
    """A primary-HDU FITS file of arr_xy [naxis1, naxis2] written by hand (raw values stored)."""
    n1, n2 = arr_xy.shape
    cards = ["SIMPLE  = %20s" % "T", "BITPIX  = %20d" % bitpix, "NAXIS   = %20d" % 2,
             "NAXIS1  = %20d" % n1, "NAXIS2  = %20d" % n2]
    if bscale is not None:
        cards.append("BSCALE  = %20s / scale" % repr(float(bscale)))
    if bzero is not None:
        cards.append("BZERO   = %20s" % repr(float(bzero)))
    cards += list(extra_cards) + ["END"]
    hdr = "".join(c.ljust(80) for c in cards)
    hdr += " " * (-len(hdr) % 2880)
    dt = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}[bitpix]
    data = np.ascontiguousarray(arr_xy.T).astype(dt).tobytes()
    data += b"\0" * (-len(data) % 2880)
    with open(path, "wb") as f:
        f.write(hdr.encode("ascii") + data)


@pytest.mark.parametrize("bitpix", [8, 16, 32, 64, -32, -64])
def test_fits_bitpix_scaling_and_keywords(tmp_path, bitpix):
    """Every BITPIX, BSCALE / BZERO, the reference's PXSIZE / PXNOISE keywords, a header longer
    than one block; a non-square image distinguishes the axes."""
    rng = np.random.default_rng(bitpix + 100)
    raw = rng.integers(0, 120, size=(7, 5)).astype(np.float64)
    if bitpix < 0:
        raw = raw / 8.0
    comments = ["COMMENT   padding card %d" % k for k in range(40)]
    keys = ["PXSIZE  = %20s" % "0.049", "PXNOISE = %20s / noise" % "0.0125"]
    path = str(tmp_path / "t.fits")
    _write_fits(path, raw, bitpix, comments + keys, bscale=2.0, bzero=-3.0)
    pix, pxs, pxn = qlx.read_fits(path)
    ref, _ = O.read_fits_np(path)
    assert np.array_equal(pix, raw * 2.0 - 3.0) and np.array_equal(pix, ref)
    assert pxs == 0.049 and pxn == 0.0125


def test_fits_errors(tmp_path):
    p = str(tmp_path / "x.fits")
    with open(p, "wb") as f:
        f.write(b"not a fits file" * 300)
    with pytest.raises(qlx.QlxError):
        qlx.read_fits(p)
    _write_fits(p, np.ones((4, 4)), -64)
    with open(p, "r+b") as f:
        f.truncate(2880 + 16)  # the data cut short
    with pytest.raises(qlx.QlxError):
        qlx.read_fits(p)


def test_pixel_grid_matches_the_reference_construction():
    for n, s in ((81, 0.049), (64, 0.05), (7, 0.1)):
        xv, yv = qlx.pixel_grid(n, n + 3, s)
        rx, ry = O.centred_grid(n, n + 3, s)
        assert np.array_equal(xv, rx) and np.array_equal(yv, ry)
        assert abs(xv[0] + 0.5 * n * s) == 0 and abs(xv[-1] - 0.5 * n * s) < 1e-12


def test_psf_from_shipped_fits():
    img, _ = O.read_fits_np(os.path.join(FITS, "hst_psf.fits"))
    for th in (1e-2, 1e-3, 0.05):
        psf = qlx.psf_from_image(img, th)
        ref = O.psf_crop(img, th)
        assert psf.shape == ref.shape and np.array_equal(psf, ref), th
        assert psf.shape[0] % 2 == 1 and psf.shape[1] % 2 == 1
        assert abs(psf.sum() - 1.0) < 1e-14
        # the window is as wide as the above-threshold box rounded up to odd, centred on the
        # input's middle (not on the box: an off-centre box loses its far edge, as in the reference)
        ii, jj = np.nonzero(img > th * img.max())
        bx = max(ii.max(), 18) - min(ii.min(), 18) + 1
        by = max(jj.max(), 18) - min(jj.min(), 18) + 1
        assert psf.shape == (2 * (bx // 2) + 1, 2 * (by // 2) + 1)


def test_psf_window_closed_form():
    # a 9 x 7 image whose above-threshold pixels span i = 3..5, j = 2..4 around the middle (4, 3):
    # box widths 3 -> half widths 1 -> a 3 x 3 window = the box, normalised
    img = np.full((9, 7), 1e-4)
    img[3:6, 2:5] = [[1, 2, 1], [2, 8, 2], [1, 2, 1]]
    psf = qlx.psf_from_image(img, 0.05)
    assert np.array_equal(psf, img[3:6, 2:5] / 20.0)
    # an even box width (i = 2..5: 4 pixels) gives half width 2: a 5-wide window in x
    img[2, 3] = 1.0
    assert qlx.psf_from_image(img, 0.05).shape == (5, 3)
    # an even-sized input whose box spans it (i = 0..7 of 8, middle 4, half width 4) asks for a
    # window past the far edge (i = 8): refused
    img = np.full((8, 7), 1e-4)
    img[0, 3] = img[7, 3] = img[4, 3] = 1.0
    with pytest.raises(qlx.QlxError):
        qlx.psf_from_image(img, 0.05)


def test_mask_from_image():
    m = np.zeros((5, 3))
    m[1:4, 1] = 1.0
    m[0, 0] = -2.0  # nonzero is in the mask
    full = qlx.mask_from_image(m, 5, 3)
    assert np.array_equal(full, (m != 0).astype(np.uint8))
    # a smaller mask sits centred; pixels it does not cover keep their state
    prev = np.zeros((9, 7), dtype=np.uint8)
    prev[0, 0] = 1
    got = qlx.mask_from_image(m, 9, 7, prev)
    want = prev.copy()
    want[2:7, 2:5] = (m != 0)
    assert np.array_equal(got, want)
    with pytest.raises(qlx.QlxError):
        qlx.mask_from_image(np.ones((10, 3)), 9, 7)


def test_shipped_mask_and_image_orientation_agree():
    """The shipped mask selects the lensed arcs of the shipped image: the in-mask mean is far above
    the out-of-mask mean, and transposing the mask would not do that as well (the image is not
    symmetric under x <-> y)."""
    img, _, _ = qlx.read_fits(os.path.join(FITS, "sim_hst_image.fits"))
    m, _, _ = qlx.read_fits(os.path.join(FITS, "mask.fits"))
    mask = qlx.mask_from_image(m, 81, 81).astype(bool)
    assert 0 < mask.sum() < 81 * 81
    assert not np.array_equal(img, img.T)
    assert img[mask].mean() > 10 * abs(img[~mask].mean())
