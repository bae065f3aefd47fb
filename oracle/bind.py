"""ctypes access to the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg;
never by the library.  libqtrk_oracle.so restates the reference CPU tracker's
arithmetic; libqtrk_oracle64.so is the same source with every float widened to
double (the FP64 oracle).  oracle/_ref/libqtrk_ref.so (built only where the
reference sources are mounted) wraps the reference's own standalone headers.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class OConfig(C.Structure):
    """Byte layout of QTrkComputedConfig (cputrack/qtrk_c_api.h:59-138)."""
    _pack_ = 1
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("numThreads", C.c_int), ("cuda_device", C.c_int),
        ("com_bgcorrection", C.c_float), ("zlut_minradius", C.c_float), ("zlut_radial_coverage", C.c_float),
        ("zlut_angular_coverage", C.c_float), ("zlut_roi_coverage", C.c_float), ("qi_iterations", C.c_int),
        ("qi_minradius", C.c_float), ("qi_radial_coverage", C.c_float), ("qi_angular_coverage", C.c_float),
        ("qi_roi_coverage", C.c_float), ("qi_angstep_factor", C.c_float), ("xc1_profileLength", C.c_int),
        ("xc1_profileWidth", C.c_int), ("xc1_iterations", C.c_int), ("gauss2D_iterations", C.c_int),
        ("gauss2D_sigma", C.c_float), ("downsample", C.c_int), ("zlut_radialsteps", C.c_int),
        ("zlut_angularsteps", C.c_int), ("zlut_maxradius", C.c_float), ("qi_radialsteps", C.c_int),
        ("qi_angstepspq", C.c_int), ("qi_maxradius", C.c_float),
    ]


JOB = np.dtype([("frame", "<u4"), ("timestamp", "<u4"), ("zlutIndex", "<i4"),
                ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4")])
RESULT = np.dtype([("frame", "<u4"), ("timestamp", "<u4"), ("zlutIndex", "<i4"),
                   ("ix", "<f4"), ("iy", "<f4"), ("iz", "<f4"),
                   ("x", "<f4"), ("y", "<f4"), ("z", "<f4"),
                   ("comx", "<f4"), ("comy", "<f4"), ("error", "<u4"), ("imageMean", "<f4")])


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    def __init__(self, fp64=False):
        name = "libqtrk_oracle64.so" if fp64 else "libqtrk_oracle.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        self.fp64 = fp64
        self.fl = np.float64 if fp64 else np.float32
        self.L = L = C.CDLL(path)
        P = C.c_void_p
        fl = C.c_double if fp64 else C.c_float
        self.cfl = fl
        sig = {
            "oq_nearest_pow2": (C.c_int, [C.c_int]),
            "oq_config_update": (None, [P]),
            "oq_default_settings": (None, [P]),
            "oq_interpolate": (fl, [P, C.c_int, C.c_int, fl, fl, P]),
            "oq_radial_bin_window": (None, [C.c_int, P]),
            "oq_fft": (None, [C.c_int, C.c_int, P, P]),
            "oq_fft_leaf_order": (C.c_int, [C.c_int, P]),
            "oq_lsqfit_f": (None, [C.c_int, P, P, P, P]),
            "oq_lsqfit_d": (None, [C.c_int, P, P, P, P]),
            "oq_max_interp_qi": (fl, [P, C.c_int, P]),
            "oq_max_interp_zlut": (C.c_double, [P, C.c_int, P]),
            "oq_radial_profile": (None, [P, C.c_int, C.c_int, P, C.c_int, C.c_int, fl, fl, fl, fl, fl, C.c_int]),
            "oq_normalize_profile": (None, [P, C.c_int]),
            "oq_lut_compare": (fl, [P, P, C.c_int, C.c_int, P, P, P]),
            "oq_localize_batch": (None, [P, C.c_int, P, C.c_int, C.c_int, P, P, C.c_int, P, C.c_int, P]),
            "oq_queued_run": (C.c_double, [P, C.c_int, P, C.c_int, C.c_int, P, P, C.c_int, P, C.c_int, C.c_int, P]),
            "oq_localize_trace": (None, [P, C.c_int, P, C.c_int, C.c_int, P, P, C.c_int, P, C.c_int, P, P, P]),
            "oq_build_lut": (None, [P, P, C.c_int, C.c_int, C.c_int, P, C.c_int, P, P, P, C.c_int, C.c_int]),
            "oq_power_spectrum_2d": (None, [P, C.c_int, C.c_int]),
            "oq_build_image_lut": (None, [P, P, C.c_int, C.c_int, C.c_int, P, P, P, P, C.c_int, C.c_int, C.c_int]),
            "oq_bead_xcor_map": (None, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P]),
            "oq_find_beads": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, fl, fl, P, C.c_int]),
            "oq_fourier_radial_profile": (None, [P, C.c_int, C.c_int, P, C.c_int, C.c_int]),
            "oq_finalize_lut": (None, [P, C.c_int, C.c_int]),
            "oq_image_lut_diff": (None, [P, C.c_int, C.c_int, C.c_int, P, P]),
            "oq_bspline_weights": (None, [P, fl]),
            "oq_generate_image_from_lut": (None, [P, C.c_int, C.c_int, P, C.c_int, C.c_int, fl, fl, fl, fl, fl,
                                                  C.c_int, C.c_int]),
            "oq_fisher_ref": (None, [P, C.c_int, C.c_int, P, C.c_int, C.c_int, fl, fl, fl, fl, fl, fl, fl, fl,
                                     fl]),
            "oq_erf": (fl, [fl]),
            "oq_qm_expf": (C.c_float, [C.c_float]),
            "oq_qm_erff": (C.c_float, [C.c_float]),
            "oq_qm_check": (C.c_long, [C.c_int, C.c_float, C.c_int, P, P, P]),
            "oq_lut_compare_weighted": (fl, [P, P, C.c_int, C.c_int, fl, fl, fl, P, P]),
            "oq_zbias_apply": (C.c_float, [P, C.c_int, C.c_int, C.c_float]),
            "oq_compute_zbias_table": (None, [P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P]),
            "oq_xcor1d_img": (None, [P, C.c_int, C.c_int, fl, fl, C.c_int, C.c_int, C.c_int, P]),
            "oq_quadrant_align_img": (None, [P, C.c_int, C.c_int, fl, P, C.c_int, C.c_int, fl, fl, C.c_int, fl, fl,
                                             fl, P]),
            "oq_gaussian_spot": (None, [P, C.c_int, C.c_int, fl, fl, fl, fl, fl]),
            "oq_gauss2d_mle": (None, [P, C.c_int, C.c_int, fl, fl, fl, C.c_int, fl, P]),
            "oq_quadrant_profile_img": (None, [P, C.c_int, C.c_int, fl, C.c_int, C.c_int, C.c_int, C.c_int, fl, fl,
                                               fl, fl, P]),
            "oq_com_qi_img": (None, [P, P, fl, C.c_int, P]),
            "oq_schedule_frame": (C.c_int, [P, P, C.c_int, C.c_int, C.c_int, P, C.c_int, C.c_int, P, P, P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a

    # -- config
    def config(self, **settings):
        c = OConfig()
        self.L.oq_default_settings(C.byref(c))
        for k, v in settings.items():
            setattr(c, k, v)
        self.L.oq_config_update(C.byref(c))
        return c

    def config_from_bytes(self, raw):
        c = OConfig()
        C.memmove(C.byref(c), bytes(raw)[:C.sizeof(OConfig)], C.sizeof(OConfig))
        return c

    # -- qmath.h (the Gaussian MLE's exp and erf)
    def qm_check(self, which, amax, stride):
        """Every stride-th float of magnitude <= amax through qm_expf (which 0) or qm_erff (1):
        (count, max ulps from the exact value, max ulps from glibc's float function, worst x)."""
        a, b, w = C.c_double(), C.c_double(), C.c_float()
        n = self.L.oq_qm_check(which, amax, stride, C.byref(a), C.byref(b), C.byref(w))
        return n, a.value, b.value, w.value

    # -- primitives
    def interpolate(self, img, x, y):
        img = np.ascontiguousarray(img, self.fl)
        out = C.c_int()
        v = self.L.oq_interpolate(_p(img), img.shape[1], img.shape[0], x, y, C.byref(out))
        return v, bool(out.value)

    def bin_window(self, n):
        out = np.zeros(n, self.fl)
        self.L.oq_radial_bin_window(n, _p(out))
        return out

    def fft(self, x, inverse=False):
        x = np.ascontiguousarray(x, np.complex128 if self.fp64 else np.complex64)
        out = np.zeros_like(x)
        self.L.oq_fft(len(x), 1 if inverse else 0, _p(x), _p(out))
        return out

    def fft_leaf_order(self, n):
        perm = np.zeros(n, np.int32)
        self.L.oq_fft_leaf_order(n, _p(perm))
        return perm

    def lsqfit(self, x, y, w, double=False):
        t = np.float64 if double else self.fl
        x, y, w = (np.ascontiguousarray(a, t) for a in (x, y, w))
        out = np.zeros(4, t)
        (self.L.oq_lsqfit_d if double else self.L.oq_lsqfit_f)(len(x), _p(x), _p(y), _p(w), _p(out))
        return out

    def max_interp_qi(self, data):
        d = np.ascontiguousarray(data, self.fl)
        im = C.c_int()
        return self.L.oq_max_interp_qi(_p(d), len(d), C.byref(im)), im.value

    def max_interp_zlut(self, data):
        d = np.ascontiguousarray(data, np.float64)
        im = C.c_int()
        return self.L.oq_max_interp_zlut(_p(d), len(d), C.byref(im)), im.value

    def radial_profile(self, img, nbins, nang, minr, maxr, cx, cy, mean, normalize):
        img = np.ascontiguousarray(img, self.fl)
        out = np.zeros(nbins, self.fl)
        self.L.oq_radial_profile(_p(img), img.shape[1], img.shape[0], _p(out), nbins, nang, minr, maxr, cx, cy,
                                 mean, 1 if normalize else 0)
        return out

    def normalize_profile(self, prof):
        p = np.array(prof, self.fl)
        self.L.oq_normalize_profile(_p(p), len(p))
        return p

    def lut_compare(self, prof, lut, weights=None):
        prof = np.ascontiguousarray(prof, self.fl)
        lut = np.ascontiguousarray(lut, np.float32)
        planes, res = lut.shape
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        sc = np.zeros(planes, np.float64)
        best = C.c_int()
        z = self.L.oq_lut_compare(_p(prof), _p(lut), planes, res, None if w is None else _p(w), _p(sc),
                                  C.byref(best))
        return z, best.value, sc

    def lut_compare_weighted(self, prof, lut, minr, maxr, z_estim):
        """LUTProfileCompareAdjustedWeights: (z, best plane, scores)."""
        prof = np.ascontiguousarray(prof, self.fl)
        lut = np.ascontiguousarray(lut, np.float32)
        planes, res = lut.shape
        sc = np.zeros(planes, np.float64)
        best = C.c_int()
        z = self.L.oq_lut_compare_weighted(_p(prof), _p(lut), planes, res, minr, maxr, z_estim, _p(sc),
                                           C.byref(best))
        return z, best.value, sc

    def zbias_apply(self, row, planes, z):
        """ZLUTBiasCorrection of z with one bead's table row."""
        row = np.ascontiguousarray(row, np.float32)
        return self.L.oq_zbias_apply(_p(row), len(row), planes, z)

    def compute_zbias_table(self, cfg, luts, bias_planes, ovs=4, spline=True):
        """ComputeZBiasCorrection: [count][bias_planes] float32."""
        luts = np.ascontiguousarray(luts, np.float32)
        out = np.zeros((luts.shape[0], bias_planes), np.float32)
        self.L.oq_compute_zbias_table(C.byref(cfg), _p(luts), luts.shape[0], luts.shape[1], bias_planes, ovs,
                                      1 if spline else 0, _p(out))
        return out

    def xcor1d(self, img, cx, cy, iterations, profile_width, length):
        """ComputeXCorInterpolated from (cx, cy): (x, y, boundary)."""
        img = np.ascontiguousarray(img, self.fl)
        out = np.zeros(3, self.fl)
        self.L.oq_xcor1d_img(_p(img), img.shape[1], img.shape[0], cx, cy, iterations, profile_width, length, _p(out))
        return float(out[0]), float(out[1]), bool(out[2])

    def quadrant_align(self, img, mean, lut, minr, maxr, angsteps, x, y, z):
        """QuadrantAlign of (x, y) against the LUT row at z: (x, y, boundary)."""
        img = np.ascontiguousarray(img, self.fl)
        lut = np.ascontiguousarray(lut, np.float32)
        out = np.zeros(3, self.fl)
        self.L.oq_quadrant_align_img(_p(img), img.shape[1], img.shape[0], mean, _p(lut), lut.shape[0], lut.shape[1],
                                     minr, maxr, angsteps, x, y, z, _p(out))
        return float(out[0]), float(out[1]), bool(out[2])

    def bspline_weights(self, t):
        w = np.zeros(4, self.fl)
        self.L.oq_bspline_weights(_p(w), t)
        return w

    def generate_image_from_lut(self, w, h, lut, minr, maxr, x, y, z, spline=True, ovs=4):
        lut = np.ascontiguousarray(lut, np.float32)
        img = np.zeros((h, w), self.fl)
        self.L.oq_generate_image_from_lut(_p(img), w, h, _p(lut), lut.shape[0], lut.shape[1], minr, maxr, x, y,
                                          z, 1 if spline else 0, ovs)
        return img

    def fisher_ref(self, w, h, lut, minr, maxr, pos, delta, max_value):
        """SampleFisherMatrix::Compute: 3x3 (x, y, z) Fisher matrix by central differences."""
        lut = np.ascontiguousarray(lut, np.float32)
        F = np.zeros(9, self.fl)
        self.L.oq_fisher_ref(_p(F), w, h, _p(lut), lut.shape[0], lut.shape[1], minr, maxr, *pos, *delta, max_value)
        return F.reshape(3, 3)

    def erf(self, x):
        return self.L.oq_erf(x)

    def gaussian_spot(self, w, h, x, y, sigma, I0, bg):
        img = np.zeros((h, w), self.fl)
        self.L.oq_gaussian_spot(_p(img), w, h, x, y, sigma, I0, bg)
        return img

    def gauss2d_mle(self, img, mean, x0, y0, iters, sigma):
        img = np.ascontiguousarray(img, self.fl)
        out = np.zeros(4, self.fl)
        self.L.oq_gauss2d_mle(_p(img), img.shape[1], img.shape[0], mean, x0, y0, iters, sigma, _p(out))
        return out

    # -- whole pipeline
    def localize(self, cfg, mode, luts, rois, jobs, zcmp=None):
        """ProcessJob for every ROI in order (single thread); returns RESULT records."""
        rois = np.ascontiguousarray(rois)
        pdt = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}[rois.dtype]
        jobs = np.ascontiguousarray(jobs, JOB)
        out = np.zeros(len(jobs), RESULT)
        luts_a = None if luts is None else np.ascontiguousarray(luts, np.float32)
        zc = None if zcmp is None else np.ascontiguousarray(zcmp, np.float32)
        nl, npl = (0, 0) if luts_a is None else luts_a.shape[:2]
        self.L.oq_localize_batch(C.byref(cfg), int(mode), None if luts_a is None else _p(luts_a), nl, npl,
                                 None if zc is None else _p(zc), _p(rois), pdt, _p(jobs), len(jobs), _p(out))
        return out

    def localize_trace(self, cfg, mode, luts, rois, jobs, zcmp=None, calib=None):
        """As localize, plus each ROI's best LUT plane (-1 without Z) and radial profile."""
        rois = np.ascontiguousarray(rois)
        pdt = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}[rois.dtype]
        jobs = np.ascontiguousarray(jobs, JOB)
        out = np.zeros(len(jobs), RESULT)
        best = np.zeros(len(jobs), np.int32)
        prof = np.zeros((len(jobs), max(cfg.zlut_radialsteps, 1)), self.fl)
        luts_a = None if luts is None else np.ascontiguousarray(luts, np.float32)
        zc = None if zcmp is None else np.ascontiguousarray(zcmp, np.float32)
        nl, npl = (0, 0) if luts_a is None else luts_a.shape[:2]
        self.L.oq_localize_trace(C.byref(cfg), int(mode), None if luts_a is None else _p(luts_a), nl, npl,
                                 None if zc is None else _p(zc), _p(rois), pdt, _p(jobs), len(jobs), _p(out),
                                 _p(best), _p(prof))
        return out, best, prof

    def quadrant_profile(self, img, mean, A, angular_steps, q, nr, minr, maxr, cx, cy):
        """ComputeQuadrantProfile of quadrant q (0..3) with ComputeQI's table of A directions."""
        img = np.ascontiguousarray(img, self.fl)
        out = np.zeros(nr, self.fl)
        self.L.oq_quadrant_profile_img(_p(img), img.shape[1], img.shape[0], mean, A, angular_steps, q, nr, minr, maxr,
                                       cx, cy, _p(out))
        return out

    def com_qi(self, cfg, img, bgcorr=0.0, window=True):
        """ComputeMeanAndCOM(bgcorr) then ComputeQI with cfg's settings, with or without the radial bin
        window: (x, y, boundary hit)."""
        img = np.ascontiguousarray(img, self.fl)
        out = np.zeros(3, self.fl)
        self.L.oq_com_qi_img(C.byref(cfg), _p(img), bgcorr, int(window), _p(out))
        return float(out[0]), float(out[1]), bool(out[2])

    def schedule_frame(self, cfg, frame, positions, job):
        """QueuedTracker::ScheduleFrame: (ROIs [n][H][W] in the frame's pixel type, jobs) of the
        ROIs it queues, in order.  frame: 2-D array, any row pitch (its strides[0])."""
        frame = np.asarray(frame)
        assert frame.strides[1] == frame.itemsize
        pdt = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}[frame.dtype]
        pos = np.ascontiguousarray(positions, np.int32).reshape(-1, 2)
        job = np.ascontiguousarray(np.asarray(job, JOB).reshape(1))
        rois = np.zeros((len(pos), cfg.height, cfg.width), frame.dtype)
        jobs = np.zeros(len(pos), JOB)
        base = frame.__array_interface__["data"][0]
        n = self.L.oq_schedule_frame(C.byref(cfg), C.c_void_p(base), frame.strides[0], frame.shape[1],
                                     frame.shape[0], _p(pos), len(pos), pdt, _p(job), _p(rois), _p(jobs))
        return rois[:n], jobs[:n]

    def power_spectrum_2d(self, img):
        """FFT2D::Apply: |FFT2(img)|^2 with the zero frequency at (w/2, h/2)."""
        out = np.array(img, self.fl, order="C")
        self.L.oq_power_spectrum_2d(_p(out), out.shape[1], out.shape[0])
        return out

    def bead_xcor_map(self, img, sx, sy, roi):
        """The correlation map of BeadFinder::Find, [H][W] at the next power-of-two size."""
        im = np.ascontiguousarray(img, self.fl)
        h, w = im.shape
        W, H = 1 << max(0, (w - 1).bit_length()), 1 << max(0, (h - 1).bit_length())
        out = np.zeros((H, W), self.fl)
        self.L.oq_bead_xcor_map(_p(im), w, h, int(sx), int(sy), int(roi), _p(out))
        return out

    def find_beads(self, img, sx, sy, roi, img_distance, similarity):
        """BeadFinder::Find: (n, 2) int32 top-left corners (x, y), in the reference's order."""
        im = np.ascontiguousarray(img, self.fl)
        h, w = im.shape
        cap = 1 << 16
        xy = np.zeros((cap, 2), np.int32)
        n = self.L.oq_find_beads(_p(im), w, h, int(sx), int(sy), int(roi), img_distance, similarity, _p(xy), cap)
        return xy[:min(n, cap)].copy()

    def fourier_radial_profile(self, img, nbins, nang):
        """FourierRadialProfile of a copy of img."""
        im = np.array(img, self.fl, order="C")
        out = np.zeros(nbins, self.fl)
        self.L.oq_fourier_radial_profile(_p(im), im.shape[1], im.shape[0], _p(out), nbins, nang)
        return out

    def build_lut(self, cfg, images, plane, zlut, known_pos=None, normalize=False, calib=(None, None),
                  fourier=False):
        """Adds plane `plane` of every bead's LUT from its image, in place (float32 [count][planes][steps])."""
        images = np.ascontiguousarray(images)
        pdt = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}[images.dtype]
        kp = None if known_pos is None else np.ascontiguousarray(known_pos, np.float32)
        off = None if calib[0] is None else np.ascontiguousarray(calib[0], np.float32)
        gain = None if calib[1] is None else np.ascontiguousarray(calib[1], np.float32)
        assert zlut.dtype == np.float32 and zlut.flags.c_contiguous
        self.L.oq_build_lut(C.byref(cfg), _p(images), pdt, len(images), int(plane), None if kp is None else _p(kp),
                            1 if normalize else 0, None if off is None else _p(off),
                            None if gain is None else _p(gain), _p(zlut), zlut.shape[1], 1 if fourier else 0)

    def build_image_lut(self, cfg, images, plane, image_lut, known_pos=None):
        """Adds plane `plane` of every bead's image LUT (float32 [beads][planes][h][w]) in place."""
        images = np.ascontiguousarray(images)
        pdt = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}[images.dtype]
        kp = None if known_pos is None else np.ascontiguousarray(known_pos, np.float32)
        assert image_lut.dtype == np.float32 and image_lut.flags.c_contiguous
        _, planes, ih, iw = image_lut.shape
        self.L.oq_build_image_lut(C.byref(cfg), _p(images), pdt, len(images), int(plane),
                                  None if kp is None else _p(kp), None, None, _p(image_lut), planes, ih, iw)

    def image_lut_diff(self, lut):
        """FinalizeLUT's (dz, dz2) of an image LUT [beads][planes][h][w] (float32)."""
        lut = np.ascontiguousarray(lut, np.float32)
        dz, dz2 = np.zeros_like(lut), np.zeros_like(lut)
        b, p = lut.shape[:2]
        self.L.oq_image_lut_diff(_p(lut), b, p, int(np.prod(lut.shape[2:])), _p(dz), _p(dz2))
        return dz, dz2

    def finalize_lut(self, zlut):
        self.L.oq_finalize_lut(_p(zlut), zlut.shape[0] * zlut.shape[1], zlut.shape[2])

    def queued_run(self, cfg, mode, luts, rois, jobs, nthreads, zcmp=None):
        """The QueuedCPUTracker restatement with `nthreads` workers; returns (seconds, results)."""
        rois = np.ascontiguousarray(rois)
        pdt = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}[rois.dtype]
        jobs = np.ascontiguousarray(jobs, JOB)
        out = np.zeros(len(jobs), RESULT)
        luts_a = np.ascontiguousarray(luts, np.float32)
        zc = None if zcmp is None else np.ascontiguousarray(zcmp, np.float32)
        t = self.L.oq_queued_run(C.byref(cfg), int(mode), _p(luts_a), luts_a.shape[0], luts_a.shape[1],
                                 None if zc is None else _p(zc), _p(rois), pdt, _p(jobs), len(jobs), int(nthreads),
                                 _p(out))
        return t, out


class RefHeaders:
    """The reference's LsqQuadraticFit.h / CubicBSpline.h compiled unmodified (oracle/_ref)."""

    PATH = os.path.join(HERE, "_ref", "libqtrk_ref.so")

    def __init__(self):
        if not os.path.exists(self.PATH):
            raise RuntimeError("oracle/_ref not built (needs the reference sources)")
        self.L = L = C.CDLL(self.PATH)
        P = C.c_void_p
        for k, r, a in [("ref_max_interp_f7", C.c_float, [P, C.c_int, P]),
                        ("ref_max_interp_d7", C.c_double, [P, C.c_int, P]),
                        ("ref_lsqfit_f", None, [C.c_int, P, P, P, P]),
                        ("ref_lsqfit_d", None, [C.c_int, P, P, P, P]),
                        ("ref_bspline_weights", None, [C.c_float, P]),
                        ("ref_bspline_derivatives", None, [C.c_float, P, P]),
                        ("ref_spline_fit_max", C.c_float, [P, C.c_int])]:
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a

    def max_interp_f7(self, data, w7):
        d = np.ascontiguousarray(data, np.float32)
        w = np.ascontiguousarray(w7, np.float32)
        return self.L.ref_max_interp_f7(_p(d), len(d), _p(w))

    def max_interp_d7(self, data, w7):
        d = np.ascontiguousarray(data, np.float64)
        w = np.ascontiguousarray(w7, np.float64)
        return self.L.ref_max_interp_d7(_p(d), len(d), _p(w))

    def lsqfit(self, x, y, w, double=False):
        t = np.float64 if double else np.float32
        x, y, w = (np.ascontiguousarray(a, t) for a in (x, y, w))
        out = np.zeros(4, t)
        (self.L.ref_lsqfit_d if double else self.L.ref_lsqfit_f)(len(x), _p(x), _p(y), _p(w), _p(out))
        return out

    def bspline_weights(self, t):
        w = np.zeros(4, np.float32)
        self.L.ref_bspline_weights(t, _p(w))
        return w
