// On-disk inputs of the inversion: the text image format, FITS images (data, mask, PSF) and the
// transformations the reference applies after reading them.  Host-only and context-free: these
// calls read files into caller buffers (two-call pattern: NULL arrays ask for the sizes), and the
// caller hands the arrays to qlx_set_image / qlx_set_psf.
//
//   text image        ImagePixelData::load_data             pixelgrid.cpp:5797-5904
//   FITS image        ImagePixelData::load_data_fits        pixelgrid.cpp:6016-6280
//   FITS mask         ImagePixelData::load_mask_fits        pixelgrid.cpp:6512-6651
//   FITS PSF          QLens::load_psf_fits                  pixelgrid.cpp:6737-6876
//
// The reference reads FITS through cfitsio (fits_read_pix into TDOUBLE, which applies BSCALE and
// BZERO); the reader here follows the FITS standard (NOST 100-2.0 / Pence et al. 2010) for a
// primary-HDU image: 2880-byte blocks of 80-character cards up to END, then big-endian pixels,
// NAXIS1 varying fastest.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/qlx.h"

namespace {

thread_local std::string g_io_err;

int io_fail(int code, const std::string& msg)
{
    g_io_err = msg;
    return code;
}

// every whitespace-separated number of a text file, as ifstream >> double reads them
bool read_numbers(const std::string& path, std::vector<double>& out)
{
    std::ifstream f(path.c_str());
    if (!f) return false;
    out.clear();
    double v;
    while (f >> v) out.push_back(v);
    return true;
}

struct FitsImage {
    int naxis = 0;
    long long n1 = 1, n2 = 1;
    int bitpix = 0;
    double bscale = 1.0, bzero = 0.0;
    double pxsize = NAN, pxnoise = NAN;
    size_t data_off = 0;
};

std::string trim(const std::string& s)
{
    size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// value field of a card (columns 11-80) up to a '/' comment; string values are not needed here
std::string card_value(const char* card)
{
    std::string v(card + 10, 70);
    size_t sl = v.find('/');
    if (sl != std::string::npos) v = v.substr(0, sl);
    return trim(v);
}

int parse_header(const std::vector<unsigned char>& buf, FitsImage& im)
{
    if (buf.size() < 2880) return io_fail(QLX_ERR_IO, "not a FITS file (shorter than one 2880-byte block)");
    if (memcmp(buf.data(), "SIMPLE  =", 9) != 0) return io_fail(QLX_ERR_IO, "not a FITS file (no SIMPLE card)");
    bool end = false, have_bitpix = false, have_naxis = false;
    size_t pos = 0;
    for (; pos + 80 <= buf.size() && !end; pos += 80) {
        const char* card = (const char*)buf.data() + pos;
        const std::string key = trim(std::string(card, 8));
        if (key == "END") {
            end = true;
            continue;
        }
        if (card[8] != '=') continue;  // COMMENT, HISTORY, blank
        const std::string val = card_value(card);
        char* e = nullptr;
        if (key == "BITPIX") {
            im.bitpix = (int)strtol(val.c_str(), &e, 10);
            have_bitpix = true;
        } else if (key == "NAXIS") {
            im.naxis = (int)strtol(val.c_str(), &e, 10);
            have_naxis = true;
        } else if (key == "NAXIS1") {
            im.n1 = strtoll(val.c_str(), &e, 10);
        } else if (key == "NAXIS2") {
            im.n2 = strtoll(val.c_str(), &e, 10);
        } else if (key == "BSCALE") {
            im.bscale = strtod(val.c_str(), &e);
        } else if (key == "BZERO") {
            im.bzero = strtod(val.c_str(), &e);
        } else if (key == "PXSIZE") {  // the reference's own keywords (pixelgrid.cpp:6094-6110)
            im.pxsize = strtod(val.c_str(), &e);
        } else if (key == "PXNOISE") {
            im.pxnoise = strtod(val.c_str(), &e);
        }
    }
    if (!end) return io_fail(QLX_ERR_IO, "FITS header has no END card");
    if (!have_bitpix || !have_naxis) return io_fail(QLX_ERR_IO, "FITS header lacks BITPIX or NAXIS");
    if (im.naxis < 1) return io_fail(QLX_ERR_IO, "only 1D or 2D FITS images are supported (NAXIS = 0)");
    if (im.naxis < 2) im.n2 = 1;
    const int b = im.bitpix;
    if (b != 8 && b != 16 && b != 32 && b != 64 && b != -32 && b != -64) return io_fail(QLX_ERR_IO, "unsupported FITS BITPIX");
    if (im.n1 < 1 || im.n2 < 1 || im.n1 > 65535 || im.n2 > 65535) return io_fail(QLX_ERR_RANGE, "FITS image axes outside 1..65535");
    im.data_off = (pos + 2879) / 2880 * 2880;
    const size_t need = im.data_off + (size_t)(im.n1 * im.n2) * (size_t)(std::abs(b) / 8);
    if (buf.size() < need) return io_fail(QLX_ERR_IO, "FITS file shorter than its image");
    return QLX_OK;
}

double be_value(const unsigned char* p, int bitpix)
{
    unsigned long long u = 0;
    const int nb = std::abs(bitpix) / 8;
    for (int k = 0; k < nb; k++) u = (u << 8) | p[k];
    switch (bitpix) {
    case 8: return (double)(unsigned char)u;
    case 16: return (double)(short)(unsigned short)u;
    case 32: return (double)(int)(unsigned)u;
    case 64: return (double)(long long)u;
    case -32: {
        const unsigned w = (unsigned)u;
        float f;
        memcpy(&f, &w, 4);
        return (double)f;
    }
    default: {
        double d;
        memcpy(&d, &u, 8);
        return d;
    }
    }
}

}  // namespace

extern "C" {

const char* qlx_io_last_error(void) { return g_io_err.c_str(); }

int qlx_read_image_text(const char* root, int* nx, int* ny, double* xvals, double* yvals, double* sb, size_t sb_cap)
{
    if (!root || !nx || !ny) return io_fail(QLX_ERR_ARG, "qlx_read_image_text: null argument");
    const std::string r(root);
    std::vector<double> xv, yv;
    if (!read_numbers(r + ".x", xv) || !read_numbers(r + ".y", yv))
        return io_fail(QLX_ERR_IO, "cannot read " + r + ".x or " + r + ".y");
    if (xv.size() < 2 || yv.size() < 2) return io_fail(QLX_ERR_IO, "pixel corner files need at least two values");
    if (xv.size() - 1 > 65535 || yv.size() - 1 > 65535) return io_fail(QLX_ERR_RANGE, "image wider or taller than 65535 pixels");
    *nx = (int)xv.size() - 1;  // npixels = values - 1 (pixelgrid.cpp:5846-5856)
    *ny = (int)yv.size() - 1;
    const size_t npix = (size_t)*nx * *ny;
    if (!xvals && !yvals && !sb) return QLX_OK;
    if (xvals) memcpy(xvals, xv.data(), xv.size() * 8);
    if (yvals) memcpy(yvals, yv.data(), yv.size() * 8);
    if (sb) {
        if (sb_cap < npix) return io_fail(QLX_ERR_ARG, "qlx_read_image_text: sb buffer smaller than nx*ny");
        std::vector<double> d;
        if (!read_numbers(r + ".dat", d)) return io_fail(QLX_ERR_IO, "cannot read " + r + ".dat");
        if (d.size() < npix) return io_fail(QLX_ERR_IO, r + ".dat holds fewer than nx*ny values");
        memcpy(sb, d.data(), npix * 8);  // x-major: i outer, j inner (pixelgrid.cpp:5887-5899)
    }
    return QLX_OK;
}

int qlx_read_fits(const char* path, int* naxis1, int* naxis2, double* pix, size_t cap, double* pxsize, double* pxnoise)
{
    if (!path || !naxis1 || !naxis2) return io_fail(QLX_ERR_ARG, "qlx_read_fits: null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return io_fail(QLX_ERR_IO, std::string("cannot open ") + path);
    std::vector<unsigned char> buf;
    unsigned char chunk[65536];
    size_t n;
    while ((n = fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + n);
    fclose(f);
    FitsImage im;
    int rc = parse_header(buf, im);
    if (rc) return rc;
    *naxis1 = (int)im.n1;
    *naxis2 = (int)im.n2;
    if (pxsize) *pxsize = im.pxsize;
    if (pxnoise) *pxnoise = im.pxnoise;
    if (!pix) return QLX_OK;
    const size_t npix = (size_t)(im.n1 * im.n2);
    if (cap < npix) return io_fail(QLX_ERR_ARG, "qlx_read_fits: pixel buffer smaller than naxis1*naxis2");
    const int nb = std::abs(im.bitpix) / 8;
    const unsigned char* d = buf.data() + im.data_off;
    // row j = fpixel[1]-1 holds naxis1 values along x; stored as pix[i*naxis2 + j] like
    // surface_brightness[i][j] (pixelgrid.cpp:6241-6249)
    for (long long j = 0; j < im.n2; j++)
        for (long long i = 0; i < im.n1; i++) {
            const double raw = be_value(d + (size_t)(j * im.n1 + i) * nb, im.bitpix);
            pix[(size_t)i * im.n2 + j] = (im.bscale == 1.0 && im.bzero == 0.0) ? raw : raw * im.bscale + im.bzero;
        }
    return QLX_OK;
}

int qlx_pixel_grid(int nx, int ny, double pixel_size, double* xvals, double* yvals)
{
    if (nx < 1 || ny < 1 || !(pixel_size > 0) || !xvals || !yvals) return io_fail(QLX_ERR_ARG, "qlx_pixel_grid: bad arguments");
    // load_data_fits with a pixel size (pixelgrid.cpp:6198-6213)
    const double xstep = pixel_size, ystep = pixel_size;
    const double xmax = 0.5 * nx * pixel_size, ymax = 0.5 * ny * pixel_size;
    const double xmin = -xmax, ymin = -ymax;
    double x = xmin, y = ymin;
    for (int i = 0; i <= nx; i++, x += xstep) xvals[i] = x;
    for (int j = 0; j <= ny; j++, y += ystep) yvals[j] = y;
    return QLX_OK;
}

int qlx_psf_from_image(const double* img, int nx, int ny, double threshold, int* pnx, int* pny, double* psf, size_t cap)
{
    if (!img || nx < 1 || ny < 1 || !pnx || !pny) return io_fail(QLX_ERR_ARG, "qlx_psf_from_image: bad arguments");
    // load_psf_fits after the read (pixelgrid.cpp:6790-6840): the box of pixels above
    // threshold x peak, a window of 2 half + 1 pixels centred on (nx/2, ny/2), unit sum
    double peak = -1e30;
    for (size_t k = 0; k < (size_t)nx * ny; k++)
        if (img[k] > peak) peak = img[k];
    const int imid = nx / 2, jmid = ny / 2;
    int imin = imid, imax = imid, jmin = jmid, jmax = jmid;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            if (img[(size_t)i * ny + j] > threshold * peak) {
                if (i < imin) imin = i;
                if (i > imax) imax = i;
                if (j < jmin) jmin = j;
                if (j > jmax) jmax = j;
            }
    const int nxh = (imax - imin + 1) / 2, nyh = (jmax - jmin + 1) / 2;
    *pnx = 2 * nxh + 1;
    *pny = 2 * nyh + 1;
    // the window must lie inside the input (the reference reads outside it otherwise)
    if (imid - nxh < 0 || imid + nxh >= nx || jmid - nyh < 0 || jmid + nyh >= ny)
        return io_fail(QLX_ERR_RANGE, "PSF window above the threshold is off-centre: it reaches past the input image");
    if (!psf) return QLX_OK;
    if (cap < (size_t)*pnx * *pny) return io_fail(QLX_ERR_ARG, "qlx_psf_from_image: PSF buffer too small");
    for (int ii = 0, i = imid - nxh; ii < *pnx; i++, ii++)
        for (int jj = 0, j = jmid - nyh; jj < *pny; j++, jj++) psf[(size_t)ii * *pny + jj] = img[(size_t)i * ny + j];
    double norm = 0;
    for (int i = 0; i < *pnx; i++)
        for (int j = 0; j < *pny; j++) norm += psf[(size_t)i * *pny + j];
    for (int i = 0; i < *pnx; i++)
        for (int j = 0; j < *pny; j++) psf[(size_t)i * *pny + j] /= norm;
    return QLX_OK;
}

int qlx_mask_from_image(const double* m, int mnx, int mny, int nx, int ny, unsigned char* mask)
{
    if (!m || !mask || mnx < 1 || mny < 1 || nx < 1 || ny < 1) return io_fail(QLX_ERR_ARG, "qlx_mask_from_image: bad arguments");
    // load_mask_fits into the primary mask (pixelgrid.cpp:6532-6620): a mask smaller than the
    // image along both axes is centred with offsets ((nx - mnx)/2, (ny - mny)/2) and the pixels it
    // does not cover keep their state; a larger one is refused
    int ox = 0, oy = 0;
    if (mnx != nx || mny != ny) {
        if (mnx < nx && mny < ny) {
            ox = (nx - mnx) / 2;
            oy = (ny - mny) / 2;
        } else {
            return io_fail(QLX_ERR_ARG, "mask has more pixels than the image along x and/or y");
        }
    }
    for (int j = 0; j < mny; j++)
        for (int i = 0; i < mnx; i++) mask[(size_t)(i + ox) * ny + (j + oy)] = (m[(size_t)i * mny + j] == 0.0) ? 0 : 1;
    return QLX_OK;
}

}  // extern "C"
