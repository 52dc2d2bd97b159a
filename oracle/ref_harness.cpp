// Driver for the compiled QLens reference (built by oracle/ref.mk into oracle/_ref/).
//
// Checker only: this program links the reference's own objects and is never part of the
// product.  It sets up a QLens object from a command script (the reference's own command
// language, commands.cpp), then runs one of these modes and writes raw little-endian
// arrays plus a text file of scalars into <outdir>:
//
//   invert  <script> <outdir> [sigma] [nrep] [psf=<file>:<nx>:<ny>] [mask=<file>:<nx>:<ny>]
//       Runs QLens::invert_surface_brightness_map_from_data (lens.cpp:12246) on the image
//       loaded by the script, with a uniform noise map of dispersion sigma (the state that
//       'sbmap generate_uniform_noisemap' produces, commands.cpp:9453-9458).  Dumps -2 ln E,
//       chisq0, ln det F, ln det R, active counts, the subpixel source points (rays), the
//       image mapping flags, the model image and the source-cell amplitudes.  With nrep > 1
//       the inversion is repeated and the median wall time per call over all but the first
//       (warm-up) call is reported.  With
//       psf=..., the PSF matrix is injected from a raw x-major file (QLens's psf_matrix with
//       use_input_psf_matrix, qlens.h:877-880), as the reference does for a loaded PSF.  With
//       mask=..., primary mask 0 of the image is set from a raw x-major byte file (the state
//       'sbmap loadmask' leaves, pixelgrid.cpp:6512-6647; FITS loading is compiled out here).
//   deflect <script> <outdir> points=<points.bin>:<n>
//       Evaluates QLens::deflection (lenscalc.cpp:81-113) and the per-lens
//       LensProfile::deflection / hessian (profile.cpp:1782-1805) at n points (x,y pairs).
//   patterson <script> <outdir>
//       Dumps the Gauss-Patterson nodes and weights held by lens 0 (gauss.cpp:521-2099).
//   mock <script> <outdir> [psf=<file>:<nx>:<ny>]
//       Runs QLens::plot_lensed_surface_brightness (lens.cpp:11882-12066) on the no-data image
//       grid the script sets up ('grid xmin xmax ymin ymax', 'img_npixels'): the analytic
//       sources of the script ray-traced per subpixel and averaged
//       (ImagePixelGrid::find_surface_brightness, pixelgrid.cpp:11700-11940), PSF-convolved
//       (PSF_convolution_pixel_vector, pixelgrid.cpp:14224-14496) and, with
//       'simulate_pixel_noise on', given noise from the generator seeded by 'random_seed'
//       (add_pixel_noise, pixelgrid.cpp:10987-10996).  Dumps the image x-major (mock.bin).
//   rng <script> <outdir> <seed> <n>
//       The first n normal deviates of the reference's generator (Random, rand.h:19-76) after
//       set_random_seed(seed) (deviates.bin), and the first n raw 64-bit outputs of a second
//       generator with the same seed (int64.bin).
//   psf <script> <outdir> <xstep> <ystep>
//       QLens::generate_PSF_matrix (pixelgrid.cpp:14601-14668) for the script's psf_width and
//       psf_threshold on pixels of xstep by ystep (psf.bin x-major, sizes in scalars.txt).
//
// Build: make -f oracle/ref.mk (needs /root/reference; output lives only in oracle/_ref/).

#include <omp.h>
#include <algorithm>
#include <chrono>
#include <complex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>
#include <functional>
#include <unistd.h>
// The reference keeps grid state private; this checker translation unit alone is compiled
// with -fno-access-control (oracle/ref.mk) so it can read that state after an inversion.
#include "qlens.h"
#include "pixelgrid.h"
#include "profile.h"

using namespace std;

static void write_bin(const string& path, const void* data, size_t bytes)
{
	FILE* f = fopen(path.c_str(), "wb");
	if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(2); }
	if (bytes) fwrite(data, 1, bytes, f);
	fclose(f);
}

static void install_uniform_noise(QLens& lens, double sigma)
{
	// Same end state as 'sbmap generate_uniform_noisemap' when the maps exist
	// (pixelgrid.h:569-580): noise_map = sigma, covinv_map = 1/sigma^2, use_noise_map on.
	ImagePixelData* d = lens.image_pixel_data;
	if (!d) { fprintf(stderr, "no image data loaded\n"); exit(2); }
	int nx = d->npixels_x, ny = d->npixels_y;
	if (d->noise_map == NULL) {
		d->noise_map = new double*[nx];
		for (int i = 0; i < nx; i++) d->noise_map[i] = new double[ny];
	}
	if (d->covinv_map == NULL) {
		d->covinv_map = new double*[nx];
		for (int i = 0; i < nx; i++) d->covinv_map[i] = new double[ny];
	}
	d->set_uniform_pixel_noise(sigma);
	lens.use_noise_map = true;
	lens.background_pixel_noise = sigma;
}

static int mode_invert(QLens& lens, const string& out, double sigma, int nrep)
{
	install_uniform_noise(lens, sigma);
	double chisq0 = 0, m2lnE = 0;
	vector<double> times;
	for (int r = 0; r < nrep; r++) {
		auto t0 = chrono::steady_clock::now();
		m2lnE = lens.invert_surface_brightness_map_from_data(chisq0, false);
		auto t1 = chrono::steady_clock::now();
		times.push_back(chrono::duration<double>(t1 - t0).count());
	}
	// with more than one repetition the first is a warm-up and the median is over the others
	vector<double> sorted(times.begin() + (nrep > 1 ? 1 : 0), times.end());
	sort(sorted.begin(), sorted.end());
	double median = sorted.size() % 2 ? sorted[sorted.size() / 2] : 0.5 * (sorted[sorted.size() / 2 - 1] + sorted[sorted.size() / 2]);

	ImagePixelGrid* g = lens.image_pixel_grids[0];
	int nx = g->x_N, ny = g->y_N;
	// split_imgpixels off: one ray per pixel, the pixel-centre ray (center_sourcepts)
	int nsub = lens.split_imgpixels ? INTSQR(lens.default_imgpixel_nsplit) : 1;

	// rays: beta for every subpixel of every pixel, [i][j][k][2], i-major
	vector<double> rays((size_t)nx * ny * nsub * 2);
	vector<unsigned char> mapped((size_t)nx * ny);
	vector<double> model((size_t)nx * ny);
	for (int i = 0; i < nx; i++)
		for (int j = 0; j < ny; j++) {
			size_t p = (size_t)i * ny + j;
			mapped[p] = g->maps_to_source_pixel[i][j] ? 1 : 0;
			model[p] = g->surface_brightness[i][j];
			for (int k = 0; k < nsub; k++) {
				const lensvector& b = lens.split_imgpixels ? g->subpixel_center_sourcepts[i][j][k] : g->center_sourcepts[i][j];
				rays[(p * nsub + k) * 2 + 0] = b[0];
				rays[(p * nsub + k) * 2 + 1] = b[1];
			}
		}
	write_bin(out + "/rays.bin", rays.data(), rays.size() * sizeof(double));
	write_bin(out + "/mapped.bin", mapped.data(), mapped.size());
	write_bin(out + "/model.bin", model.data(), model.size() * sizeof(double));

	SourcePixelGrid* sg = g->cartesian_srcgrid;
	int snx = sg->u_N, sny = sg->w_N;
	vector<double> src((size_t)snx * sny);
	vector<int> active((size_t)snx * sny);
	for (int i = 0; i < snx; i++)
		for (int j = 0; j < sny; j++) {
			SourcePixel* c = sg->cell[i][j];
			src[(size_t)i * sny + j] = c->surface_brightness;
			active[(size_t)i * sny + j] = c->active_pixel ? c->active_index : -1;
		}
	write_bin(out + "/src.bin", src.data(), src.size() * sizeof(double));
	write_bin(out + "/active.bin", active.data(), active.size() * sizeof(int));

	FILE* f = fopen((out + "/scalars.txt").c_str(), "w");
	fprintf(f, "minus2lnE %.17g\n", m2lnE);
	fprintf(f, "chisq0 %.17g\n", chisq0);
	fprintf(f, "logdetF %.17g\n", lens.Fmatrix_log_determinant);
	fprintf(f, "logdetR %.17g\n", lens.Rmatrix_log_determinant);
	fprintf(f, "n_img_active %d\n", lens.image_npixels);
	fprintf(f, "n_src_active %d\n", lens.source_npixels);
	fprintf(f, "img_nx %d\nimg_ny %d\nnsub %d\n", nx, ny, nsub);
	fprintf(f, "src_nx %d\nsrc_ny %d\n", snx, sny);
	fprintf(f, "src_xmin %.17g\nsrc_xmax %.17g\nsrc_ymin %.17g\nsrc_ymax %.17g\n",
		sg->srcgrid_xmin, sg->srcgrid_xmax, sg->srcgrid_ymin, sg->srcgrid_ymax);
	fprintf(f, "regparam %.17g\n", sg->regparam);
	fprintf(f, "psf_nx %d\npsf_ny %d\n", lens.psf_npixels_x, lens.psf_npixels_y);
	fprintf(f, "median_seconds %.9g\n", median);
	fprintf(f, "min_seconds %.9g\nmax_seconds %.9g\n", sorted.front(), sorted.back());
	fprintf(f, "nrep %d\n", nrep);
	fprintf(f, "omp_threads %d\n", omp_get_max_threads());
	fclose(f);
	if (lens.psf_matrix) {
		vector<double> psf((size_t)lens.psf_npixels_x * lens.psf_npixels_y);
		for (int i = 0; i < lens.psf_npixels_x; i++)
			for (int j = 0; j < lens.psf_npixels_y; j++) psf[(size_t)i * lens.psf_npixels_y + j] = lens.psf_matrix[i][j];
		write_bin(out + "/psf_used.bin", psf.data(), psf.size() * sizeof(double));
	}
	return 0;
}

static int mode_deflect(QLens& lens, const string& out, const string& ptsfile, int n)
{
	vector<double> pts((size_t)2 * n);
	FILE* f = fopen(ptsfile.c_str(), "rb");
	if (!f || fread(pts.data(), sizeof(double), pts.size(), f) != pts.size()) { fprintf(stderr, "bad points file\n"); return 2; }
	fclose(f);
	int nl = lens.nlens;
	vector<double> tot((size_t)2 * n), per((size_t)2 * n * nl), hess((size_t)4 * n * nl);
	for (int p = 0; p < n; p++) {
		lensvector d;
		lens.deflection(pts[2 * p], pts[2 * p + 1], d, 0, lens.reference_zfactors, lens.default_zsrc_beta_factors);
		tot[2 * p] = d[0];
		tot[2 * p + 1] = d[1];
		for (int l = 0; l < nl; l++) {
			lensvector dl;
			lens.lens_list[l]->deflection(pts[2 * p], pts[2 * p + 1], dl);
			per[((size_t)l * n + p) * 2 + 0] = dl[0];
			per[((size_t)l * n + p) * 2 + 1] = dl[1];
			lensmatrix h;
			lens.lens_list[l]->hessian(pts[2 * p], pts[2 * p + 1], h);
			hess[((size_t)l * n + p) * 4 + 0] = h[0][0];
			hess[((size_t)l * n + p) * 4 + 1] = h[0][1];
			hess[((size_t)l * n + p) * 4 + 2] = h[1][0];
			hess[((size_t)l * n + p) * 4 + 3] = h[1][1];
		}
	}
	write_bin(out + "/def_total.bin", tot.data(), tot.size() * sizeof(double));
	write_bin(out + "/def_per_lens.bin", per.data(), per.size() * sizeof(double));
	write_bin(out + "/hess_per_lens.bin", hess.data(), hess.size() * sizeof(double));
	return 0;
}

static int mode_mock(QLens& lens, const string& out)
{
	lens.set_random_seed(lens.random_seed);
	if (!lens.plot_lensed_surface_brightness(out + "/mock_plot", false, false, false, false, false, false, false, false, false,
		false, false, false, -1, false)) {
		fprintf(stderr, "plot_lensed_surface_brightness failed\n");
		return 2;
	}
	ImagePixelGrid* g = lens.image_pixel_grids[0];
	int nx = g->x_N, ny = g->y_N;
	vector<double> img((size_t)nx * ny);
	for (int i = 0; i < nx; i++)
		for (int j = 0; j < ny; j++) img[(size_t)i * ny + j] = g->surface_brightness[i][j];
	write_bin(out + "/mock.bin", img.data(), img.size() * sizeof(double));
	FILE* f = fopen((out + "/scalars.txt").c_str(), "w");
	fprintf(f, "img_nx %d\nimg_ny %d\nrandom_seed %lld\nsigma %.17g\n", nx, ny, (long long)lens.random_seed,
		lens.background_pixel_noise);
	fprintf(f, "psf_nx %d\npsf_ny %d\n", lens.psf_npixels_x, lens.psf_npixels_y);
	fclose(f);
	return 0;
}

static int mode_rng(const string& out, unsigned long long seed, int n)
{
	Random r(seed), r64(seed);
	r.set_random_seed(seed);
	vector<double> dev((size_t)n);
	vector<unsigned long long> raw((size_t)n);
	for (int k = 0; k < n; k++) dev[k] = r.NormalDeviate();
	for (int k = 0; k < n; k++) raw[k] = r64.int64();
	write_bin(out + "/deviates.bin", dev.data(), dev.size() * sizeof(double));
	write_bin(out + "/int64.bin", raw.data(), raw.size() * sizeof(unsigned long long));
	return 0;
}

static int mode_psf(QLens& lens, const string& out, double xstep, double ystep)
{
	if (!lens.generate_PSF_matrix(xstep, ystep, false)) {
		fprintf(stderr, "generate_PSF_matrix failed\n");
		return 2;
	}
	vector<double> psf((size_t)lens.psf_npixels_x * lens.psf_npixels_y);
	for (int i = 0; i < lens.psf_npixels_x; i++)
		for (int j = 0; j < lens.psf_npixels_y; j++) psf[(size_t)i * lens.psf_npixels_y + j] = lens.psf_matrix[i][j];
	write_bin(out + "/psf.bin", psf.data(), psf.size() * sizeof(double));
	FILE* f = fopen((out + "/scalars.txt").c_str(), "w");
	fprintf(f, "psf_nx %d\npsf_ny %d\n", lens.psf_npixels_x, lens.psf_npixels_y);
	fclose(f);
	return 0;
}

static int mode_patterson(QLens& lens, const string& out)
{
	if (lens.nlens < 1) { fprintf(stderr, "need one lens\n"); return 2; }
	LensProfile* lp = lens.lens_list[0];
	write_bin(out + "/pat_points.bin", lp->pat_points, 511 * sizeof(double));
	vector<double> w;
	for (int level = 0; level < 9; level++)
		for (int j = 0; j < lp->pat_orders[level]; j++) w.push_back(lp->pat_weights[level][j]);
	write_bin(out + "/pat_weights.bin", w.data(), w.size() * sizeof(double));
	return 0;
}

int main(int argc, char* argv[])
{
	if (argc < 4) {
		fprintf(stderr, "usage: ref_harness invert|deflect|patterson|mock|rng|psf <script> <outdir> ...\n");
		return 2;
	}
	string mode = argv[1], script = argv[2], out = argv[3];

	// Same process-wide setup as the reference CLI (qlens.cpp:53-67).
	int nthreads = omp_get_max_threads();
	Grid::allocate_multithreaded_variables(nthreads);
	SourcePixel::allocate_multithreaded_variables(nthreads);
	DelaunayGrid::allocate_multithreaded_variables(nthreads);
	ImagePixelGrid::allocate_multithreaded_variables(nthreads);
	QLens::allocate_multithreaded_variables(nthreads);

	int rc = 0;
	{
		QLens lens;
		lens.set_verbal_mode(false);
		lens.set_inversion_nthreads(nthreads);
		lens.set_quit_after_error(true);
		lens.set_suppress_plots(true);
		if (!lens.open_script_file(script)) { fprintf(stderr, "cannot open %s\n", script.c_str()); return 2; }
		lens.set_quit_after_reading_file(true);
		lens.process_commands(true);

		// key=value options: psf=<file>:<nx>:<ny>, points=<file>:<n>, mask=<file>:<nx>:<ny>
		string psf_file, pts_file, mask_file;
		int psf_nx = 0, psf_ny = 0, npts = 0, mask_nx = 0, mask_ny = 0;
		vector<string> pos;
		for (int a = 4; a < argc; a++) {
			string arg = argv[a];
			if (arg.compare(0, 4, "psf=") == 0) {
				char buf[4096];
				if (sscanf(arg.c_str() + 4, "%4095[^:]:%d:%d", buf, &psf_nx, &psf_ny) != 3) { fprintf(stderr, "bad psf=\n"); return 2; }
				psf_file = buf;
			} else if (arg.compare(0, 7, "points=") == 0) {
				char buf[4096];
				if (sscanf(arg.c_str() + 7, "%4095[^:]:%d", buf, &npts) != 2) { fprintf(stderr, "bad points=\n"); return 2; }
				pts_file = buf;
			} else if (arg.compare(0, 5, "mask=") == 0) {
				char buf[4096];
				if (sscanf(arg.c_str() + 5, "%4095[^:]:%d:%d", buf, &mask_nx, &mask_ny) != 3) { fprintf(stderr, "bad mask=\n"); return 2; }
				mask_file = buf;
			} else pos.push_back(arg);
		}
		if (!mask_file.empty()) {
			// primary mask 0 of the loaded image, as ImagePixelData::load_mask_fits leaves it for a
			// primary mask (pixelgrid.cpp:6620-6647): in_mask from the file (x-major bytes, nonzero =
			// in), the extended mask keeping every pixel, n_mask_pixels recounted
			ImagePixelData* d = lens.image_pixel_data;
			if (!d || d->npixels_x != mask_nx || d->npixels_y != mask_ny) { fprintf(stderr, "mask size does not match the image\n"); return 2; }
			vector<unsigned char> mk((size_t)mask_nx * mask_ny);
			FILE* mf = fopen(mask_file.c_str(), "rb");
			if (!mf || fread(mk.data(), 1, mk.size(), mf) != mk.size()) { fprintf(stderr, "bad mask file\n"); return 2; }
			fclose(mf);
			int count = 0;
			for (int i = 0; i < mask_nx; i++)
				for (int j = 0; j < mask_ny; j++) {
					d->in_mask[0][i][j] = mk[(size_t)i * mask_ny + j] != 0;
					if (d->in_mask[0][i][j]) {
						d->extended_mask[0][i][j] = true;
						count++;
					}
				}
			d->n_mask_pixels[0] = count;
			d->extended_mask_n_neighbors[0] = -1;
		}
		if (!psf_file.empty()) {
			vector<double> buf((size_t)psf_nx * psf_ny);
			FILE* pf = fopen(psf_file.c_str(), "rb");
			if (!pf || fread(buf.data(), sizeof(double), buf.size(), pf) != buf.size()) { fprintf(stderr, "bad psf file\n"); return 2; }
			fclose(pf);
			lens.psf_matrix = new double*[psf_nx];
			for (int i = 0; i < psf_nx; i++) {
				lens.psf_matrix[i] = new double[psf_ny];
				for (int j = 0; j < psf_ny; j++) lens.psf_matrix[i][j] = buf[(size_t)i * psf_ny + j];
			}
			lens.psf_npixels_x = psf_nx;
			lens.psf_npixels_y = psf_ny;
			lens.use_input_psf_matrix = true;
		}
		if (mode == "invert") {
			double sigma = (pos.size() > 0) ? atof(pos[0].c_str()) : 0.01;
			int nrep = (pos.size() > 1) ? atoi(pos[1].c_str()) : 1;
			rc = mode_invert(lens, out, sigma, nrep);
		} else if (mode == "deflect") {
			if (pts_file.empty()) { fprintf(stderr, "deflect needs points=<file>:<n>\n"); return 2; }
			rc = mode_deflect(lens, out, pts_file, npts);
		} else if (mode == "patterson") {
			rc = mode_patterson(lens, out);
		} else if (mode == "mock") {
			rc = mode_mock(lens, out);
		} else if (mode == "rng") {
			if (pos.size() < 2) { fprintf(stderr, "rng needs <seed> <n>\n"); return 2; }
			rc = mode_rng(out, strtoull(pos[0].c_str(), NULL, 10), atoi(pos[1].c_str()));
		} else if (mode == "psf") {
			if (pos.size() < 2) { fprintf(stderr, "psf needs <xstep> <ystep>\n"); return 2; }
			rc = mode_psf(lens, out, atof(pos[0].c_str()), atof(pos[1].c_str()));
		} else {
			fprintf(stderr, "unknown mode %s\n", mode.c_str());
			rc = 2;
		}
		// The reference's own teardown of a QLens object after an inversion is not exercised by
		// its CLI and crashes here; every output is already on disk, so leave without it.
		fflush(stdout);
		fflush(stderr);
		_exit(rc);
	}
	return rc;
}
