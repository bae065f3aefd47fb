/*
 * qtrk_oracle.h -- CPU oracle for the bead-localization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch C restatement of the
 * reference QueuedCPUTracker/CPUTracker arithmetic (cputrack/cpu_tracker.cpp,
 * cputrack/utils.cpp, cputrack/LsqQuadraticFit.h, cputrack/kissfft.h,
 * cputrack/QueuedTracker.cpp, cputrack/QueuedCPUTracker.cpp).  It is the
 * checker the MI355X library is compared against and the timed CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the library itself never links or calls it.
 *
 * Built twice from one source: libqtrk_oracle.so (the reference's mixed
 * float/double arithmetic, operation for operation) and libqtrk_oracle64.so
 * (-DORACLE_FP64: every float of the reference widened to double; the "FP64
 * oracle" that results' errors are reported against).
 *
 * Each function cites the reference file:line it restates.  How each one is
 * pinned (which test checks it against what) is listed in the test files
 * tests/test_oracle_*.py.  Parity unpinned: libm ulps of the reference's
 * Windows build (cosf/sinf/expf there vs glibc here).
 */
#ifndef QTRK_ORACLE_H
#define QTRK_ORACLE_H

#include <stdint.h>

#ifdef ORACLE_FP64
typedef double fl;
#else
typedef float fl;
#endif

typedef struct { fl re, im; } ocpx;

/* Same byte layout as QTrkSettings / QTrkComputedConfig (cputrack/qtrk_c_api.h:59-138). */
#pragma pack(push, 1)
typedef struct {
	int width, height, numThreads, cuda_device;
	float com_bgcorrection;
	float zlut_minradius, zlut_radial_coverage, zlut_angular_coverage, zlut_roi_coverage;
	int qi_iterations;
	float qi_minradius, qi_radial_coverage, qi_angular_coverage, qi_roi_coverage, qi_angstep_factor;
	int xc1_profileLength, xc1_profileWidth, xc1_iterations;
	int gauss2D_iterations;
	float gauss2D_sigma;
	int downsample;
	int zlut_radialsteps, zlut_angularsteps;
	float zlut_maxradius;
	int qi_radialsteps, qi_angstepspq;
	float qi_maxradius;
} OConfig;

typedef struct { uint32_t frame, timestamp; int32_t zlutIndex; float ix, iy, iz; } OJob;
typedef struct { OJob job; float x, y, z; float comx, comy; uint32_t error; float imageMean; } OResult;
#pragma pack(pop)

/* kissfft-structured mixed-radix FFT plan (cputrack/kissfft.h:54-297) */
typedef struct {
	int n, inverse, nstages;
	int radix[32], rem[32];
	ocpx* tw;
} OFFT;

/* Everything a worker needs that is shared between workers. */
typedef struct {
	OConfig cfg;
	int mode;                  /* LocalizeModeEnum bits */
	const float* zluts;        /* [count][planes][zlut_radialsteps] or NULL */
	int zlut_count, zlut_planes;
	const float* zcmp;         /* radial weights [zlut_radialsteps] or NULL */
	const float* calib_offset; /* [count][h][w] or NULL */
	const float* calib_gain;
	fl* qi_window;             /* [qi_radialsteps] */
} OTracker;

/* Per-worker state of CPUTracker (cputrack/cpu_tracker.h:23-118). */
typedef struct {
	int w, h;
	fl* img;
	fl mean, stdev;
	int ndirs; fl* qdirs;      /* quadrant directions (x,y) pairs */
	OFFT fwd, inv;             /* QI FFTs of length 2*qi_radialsteps */
	OFFT xfwd, xinv;           /* XCor FFTs */
	OFFT qafwd, qainv;         /* QuadrantAlign FFTs of length 2*zlut_radialsteps */
} OWorker;

/* Optional intermediate outputs of one job (for the parity tests). */
typedef struct {
	fl comx, comy, mean, stdev;
	fl qi_x[16], qi_y[16];     /* centre after each QI iteration */
	int qi_imax_x[16], qi_imax_y[16];
	fl* profile;               /* [zlut_radialsteps] radial profile used for Z, or NULL */
	double* scores;            /* [planes], or NULL */
	int best_plane;
} OTrace;

#ifdef __cplusplus
extern "C" {
#endif

int oq_nearest_pow2(int v);
void oq_config_update(OConfig* c);
void oq_default_settings(OConfig* c);

fl oq_interpolate(const fl* img, int w, int h, fl x, fl y, int* outside);
void oq_radial_bin_window(int n, fl* out);

int oq_fft_init(OFFT* p, int n, int inverse);
void oq_fft_free(OFFT* p);
void oq_fft_transform(const OFFT* p, const ocpx* in, ocpx* out);
void oq_fft(int n, int inverse, const fl* in_ri, fl* out_ri);
int oq_fft_leaf_order(int n, int* perm);

void oq_lsqfit_f(int n, const fl* X, const fl* Y, const fl* W, fl* abcd);
void oq_lsqfit_d(int n, const double* X, const double* Y, const double* W, double* abcd);
fl oq_max_interp_qi(const fl* data, int len, int* imax);
double oq_max_interp_zlut(const double* data, int len, int* imax);

void oq_com(OWorker* wk, fl bgcorr, fl* comx, fl* comy);
void oq_quadrant_profile(OWorker* wk, fl* dst, int nr, int angsteps, int q, fl minr, fl maxr,
                         fl cx, fl cy, const fl* weights);
void oq_qi(OWorker* wk, fl cx, fl cy, int iters, int nr, int angstepspq, float factor,
           fl minr, fl maxr, const fl* weights, fl* outx, fl* outy, int* boundary, OTrace* tr);
fl oq_qi_offset(OWorker* wk, ocpx* profile, int nr, int* imax);
void oq_radial_profile(const fl* img, int w, int h, fl* dst, int nbins, int nang, fl minr, fl maxr,
                       fl cx, fl cy, fl mean, int normalize);
void oq_normalize_profile(fl* prof, int n);
fl oq_lut_compare(const fl* prof, const float* lut, int planes, int res, const float* weights,
                  double* scores, int* best);

int oq_worker_init(OWorker* wk, const OTracker* t);
void oq_worker_free(OWorker* wk);
void oq_set_image(OWorker* wk, const void* data, int pitch, int pdt);
void oq_process_job(const OTracker* t, OWorker* wk, const void* roi, int pitch, int pdt,
                    const OJob* job, OResult* out, OTrace* tr);

/* Whole-batch helpers for ctypes callers. */
void oq_localize_batch(const OConfig* cfg, int mode, const float* zluts, int zcount, int zplanes,
                       const float* zcmp, const void* rois, int pdt, const OJob* jobs, int n,
                       OResult* out);
double oq_queued_run(const OConfig* cfg, int mode, const float* zluts, int zcount, int zplanes,
                     const float* zcmp, const void* rois, int pdt, const OJob* jobs, int n,
                     int nthreads, OResult* out);

void oq_localize_trace(const OConfig* cfg, int mode, const float* zluts, int zcount, int zplanes, const float* zcmp,
                       const void* rois, int pdt, const OJob* jobs, int n, OResult* out, int* best, fl* profiles);
void oq_build_lut(const OConfig* cfg, const void* images, int pdt, int count, int plane, const float* known_pos,
                  int normalize, const float* calib_offset, const float* calib_gain, float* zlut, int planes,
                  int fourier);
/* FFT2D::Apply and FourierRadialProfile (cputrack/cpu_tracker.cpp:920-963), in place on img */
void oq_power_spectrum_2d(fl* img, int w, int h);

/* Bead finder (cputrack/BeadFinder.cpp:163-217): the correlation map (real part, [H][W] at the
 * frame's next power-of-two size) and the positions found. */
void oq_build_image_lut(const OConfig* cfg, const void* images, int pdt, int count, int plane, const float* known_pos,
                        const float* calib_offset, const float* calib_gain, float* image_lut, int planes, int ih,
                        int iw);
void oq_bead_xcor_map(const fl* img, int w, int h, int sx, int sy, int roi, fl* map);
int oq_find_beads(const fl* img, int w, int h, int sx, int sy, int roi, fl img_distance, fl similarity, int* xy,
                  int max_out);
void oq_fourier_radial_profile(fl* img, int w, int h, fl* dst, int nbins, int nang);
void oq_finalize_lut(float* zlut, int rows, int res);
/* FinalizeLUT's image LUT first and second plane differences (cputrack/QueuedCPUTracker.cpp:661-696) */
void oq_image_lut_diff(const float* lut, int beads, int planes, int npx, float* dz, float* dz2);

/* Image model and noise-free image generators (cputrack/utils.cpp:325-402, CubicBSpline.h:15-24). */
void oq_bspline_weights(fl* w, fl t);
void oq_generate_image_from_lut(fl* img, int w, int h, const float* zlut, int planes, int steps,
                                fl minr, fl maxr, fl px, fl py, fl pz, int spline, int ovs);
fl oq_erf(fl x);
float oq_qm_expf(float x);
float oq_qm_erff(float x);
long oq_qm_check(int which, float amax, int stride, double* max_exact, double* max_libc, float* worst);
void oq_gaussian_spot(fl* img, int w, int h, fl px, fl py, fl sigma, fl I0, fl bg);
/* LUTProfileCompareAdjustedWeights, QuadrantAlign, ComputeXCorInterpolated (cputrack/cpu_tracker.cpp) */
fl oq_lut_compare_weighted(const fl* prof, const float* lut, int planes, int res, fl minr, fl maxr, fl z_estim,
                           double* scores, int* best);
void oq_quadrant_align(OWorker* wk, const float* lut, int planes, int res, fl minr, fl maxr, int angsteps,
                       fl* px, fl* py, fl pz, int* boundary);
void oq_xcor1d(OWorker* wk, fl cx, fl cy, int iterations, int profileWidth, int xcorw, fl* outx, fl* outy,
               int* boundary);
void oq_xcor1d_img(const fl* img, int w, int h, fl cx, fl cy, int iterations, int profileWidth, int xcorw,
                   fl* out3);
void oq_quadrant_align_img(const fl* img, int w, int h, fl mean, const float* lut, int planes, int res, fl minr,
                           fl maxr, int angsteps, fl x, fl y, fl z, fl* out3);
/* ZLUTBiasCorrection and ComputeZBiasCorrection (cputrack/QueuedTracker.cpp:115-189); table is
 * [count][biasPlanes] */
float oq_zbias_apply(const float* row, int w, int planes, float z);
void oq_compute_zbias_table(const OConfig* cfg, const float* zluts, int count, int planes, int biasPlanes, int ovs,
                            int spline, float* table);
/* SampleFisherMatrix::Compute (cputrack/FisherMatrix.h:36-104): 3x3 (x, y, z)
 * Fisher matrix by central differences of GenerateImageFromLUT (spline, ovs 4),
 * written row-major into F[9]. */
void oq_fisher_ref(fl* F, int w, int h, const float* zlut, int planes, int steps, fl minr, fl maxr, fl px, fl py,
                   fl pz, fl dx, fl dy, fl dz, fl maxValue);
void oq_gauss2d_mle(const fl* img, int w, int h, fl mean, fl x0, fl y0, int iters, fl sigma,
                    fl* out4);
/* QueuedTracker::ScheduleFrame (cputrack/QueuedTracker.cpp:91-111): the ROIs it queues, copied
 * as QueuedCPUTracker::ScheduleLocalization copies them, and their jobs; returns the count */
/* Pieces of the path on a float image, for the oracle's own pins: the quadrant profile with
 * ComputeQI's direction table of A entries, and COM then QI with or without the bin window. */
void oq_quadrant_profile_img(const fl* img, int w, int h, fl mean, int A, int angularSteps, int q, int nr, fl minr,
                             fl maxr, fl cx, fl cy, fl* dst);
void oq_com_qi_img(const OConfig* cfg, const fl* img, fl bgcorr, int window, fl* out3);
int oq_schedule_frame(const OConfig* cfg, const void* img, int pitch, int width, int height, const int* positions,
                      int numROI, int pdt, const OJob* job, void* rois_out, OJob* jobs_out);

#ifdef __cplusplus
}
#endif
#endif
