// tracker_common.cpp -- implementation-independent parts of the tracker
// interface: config derivation, frame -> ROI extraction, Z-bias correction,
// and the utility exports of the reference's utils.h that the C# binding uses.
#include "ImageData.h"
#include "QueuedTracker.h"
#include "errors.hpp"
#include "hip_tracker.hpp"
#include "qtrk_amd.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

// cputrack/utils.cpp:644-652
int NearestPowerOf2(int v)
{
	int r = 1;
	while (r < v) r *= 2;
	if (fabsf((float)(r - v)) < fabsf((float)(r / 2 - v))) return r;
	return r / 2;
}

// QTrkComputedConfig::Update, cputrack/QueuedTracker.cpp:6-21.  The float
// arithmetic and truncations are the contract: callers derive buffer sizes
// from these integers.
void QTrkComputedConfigUpdate(QTrkComputedConfig* c)
{
	int roi = c->width / 2;
	c->zlut_maxradius = roi * c->zlut_roi_coverage;
	float zlut_perimeter = 2 * 3.141593f * c->zlut_maxradius;
	c->zlut_angularsteps = (int)(zlut_perimeter * c->zlut_angular_coverage);
	c->zlut_radialsteps = (int)((c->zlut_maxradius - c->zlut_minradius) * c->zlut_radial_coverage);
	c->qi_maxradius = roi * c->qi_roi_coverage;
	float qi_perimeter = 2 * 3.141593f * c->qi_maxradius;
	c->qi_angstepspq = (int)(qi_perimeter * c->qi_angular_coverage / 4);
	c->qi_radialsteps = (int)((c->qi_maxradius - c->qi_minradius) * c->qi_radial_coverage);
	c->qi_radialsteps = NearestPowerOf2(c->qi_radialsteps);
}

QTrkComputedConfig QTrkComputeConfig(const QTrkSettings& s)
{
	QTrkComputedConfig c;
	memset(&c, 0, sizeof(c));
	*(QTrkSettings*)&c = s;
	QTrkComputedConfigUpdate(&c);
	return c;
}

QueuedTracker::QueuedTracker() { memset(&cfg, 0, sizeof(cfg)); }
QueuedTracker::~QueuedTracker() {}

// cputrack/QueuedTracker.cpp:91-111: skip ROIs that leave the frame, address
// the rest in place, and offset the bead index by the ROI number.
int QueuedTracker::ScheduleFrame(void* imgptr, int pitch, int width, int height, ROIPosition* positions,
                                 int numROI, QTRK_PixelDataType pdt, const LocalizationJob* jobInfo)
{
	uint8_t* img = (uint8_t*)imgptr;
	int bpp = PDT_BytesPerPixel(pdt);
	int count = 0;
	for (int i = 0; i < numROI; i++) {
		const ROIPosition& pos = positions[i];
		if (pos.x < 0 || pos.y < 0 || pos.x + cfg.width > width || pos.y + cfg.height > height) continue;
		uint8_t* roiptr = &img[(size_t)pitch * pos.y + (size_t)pos.x * bpp];
		LocalizationJob job = *jobInfo;
		job.zlutIndex = i + jobInfo->zlutIndex;
		ScheduleLocalization(roiptr, pitch, pdt, &job);
		count++;
	}
	return count;
}

// Inverse bias lookup by 4 fixed-point steps, cputrack/QueuedTracker.cpp:115-140,
// with CImageData::interpolate1D (cputrack/utils.h:60-67, 99).
float QueuedTracker::ZLUTBiasCorrection(float z, int zlut_planes, int bead)
{
	if (zbias.empty() || bead < 0 || bead >= zbias_h) return z;
	const float* row = &zbias[(size_t)bead * zbias_w];
	float pos = z;
	for (int k = 0; k < 4; k++) {
		float tblpos = pos / (float)zlut_planes * zbias_w;
		// (int) as the reference's x86 build truncates: NaN and out-of-range give INT_MIN
		int fx = (tblpos == tblpos && tblpos < 2147483648.0f && tblpos >= -2147483648.0f) ? (int)tblpos : INT_MIN;
		float bias;
		if (fx < 0) bias = row[0];
		else if (fx >= zbias_w - 1) bias = row[zbias_w - 1];
		else bias = (row[fx + 1] - row[fx]) * (tblpos - fx) + row[fx];
		pos = z - bias;
	}
	return pos;
}

bool QueuedTracker::GetZLUTBiasCorrection(float* dst, int* tablePlanes, int* beads) const
{
	if (tablePlanes) *tablePlanes = zbias_w;
	if (beads) *beads = zbias_h;
	if (dst && !zbias.empty()) std::copy(zbias.begin(), zbias.end(), dst);
	return !zbias.empty();
}

void QueuedTracker::SetZLUTBiasCorrection(const float* table, int tablePlanes, int beads)
{
	if (!table || tablePlanes <= 0 || beads <= 0) {
		zbias.clear();
		zbias_w = zbias_h = 0;
		return;
	}
	zbias.assign(table, table + (size_t)tablePlanes * beads);
	zbias_w = tablePlanes;
	zbias_h = beads;
}

// The reference's non-virtual members on image types (cputrack/QueuedTracker.cpp:64-77, 143-203).
void QueuedTracker::ScheduleImageData(ImageData* data, const LocalizationJob* jobInfo)
{
	ScheduleLocalization(data->data, data->pitch(), QTrkFloat, jobInfo);
}

void QueuedTracker::ScheduleLocalization(uchar* data, int pitch, QTRK_PixelDataType pdt, uint frame, uint timestamp,
                                         vector3f* initial, uint zlutIndex)
{
	LocalizationJob j;
	memset(&j, 0, sizeof(j));
	j.frame = frame;
	j.timestamp = timestamp;
	if (initial) j.initialPos = *initial;
	j.zlutIndex = (int)zlutIndex;
	ScheduleLocalization((void*)data, pitch, pdt, &j);
}

void QueuedTracker::ComputeZBiasCorrection(int bias_planes, CImageData* result, int smpPerPixel, bool useSplineInterp)
{
	auto* t = dynamic_cast<qtrk::QueuedHIPTracker*>(this);
	if (!t) throw std::runtime_error("ComputeZBiasCorrection: not a MI355X tracker");
	int count = 0, planes = 0, steps = 0;
	GetRadialZLUTSize(count, planes, steps);
	std::vector<float> table((size_t)std::max(count, 0) * std::max(bias_planes, 0));
	t->ComputeZBiasCorrection(bias_planes, table.data(), smpPerPixel, useSplineInterp);
	if (result) *result = TImageData<float>(table.data(), bias_planes, count);
}

void QueuedTracker::SetZLUTBiasCorrection(const CImageData& data)
{
	SetZLUTBiasCorrection(data.data, data.w, data.h);
}

CImageData* QueuedTracker::GetZLUTBiasCorrection()
{
	if (zbias.empty()) return nullptr;
	CImageData* c = new CImageData(zbias_w, zbias_h);
	std::copy(zbias.begin(), zbias.end(), c->data);
	return c;
}

ImageData QueuedTracker::DebugImage(int ID)
{
	ImageData img;
	GetDebugImage(ID, &img.w, &img.h, &img.data);
	return img;
}

// ---------------------------------------------------------------------------
// Utility exports (cputrack/utils.h:182-219).  They are host-side helpers of
// the reference's DLL used by the C# GUI and by image simulation; they are
// not on the localization path.

namespace {
std::mt19937_64& rng()
{
	static thread_local std::mt19937_64 g(0x51AB0000u);
	return g;
}
}

QTRK_API void ComputeRadialProfile(float* dst, int radialSteps, int angularSteps, float minradius, float maxradius,
                                   vector2f center, ImageData* src, float mean, bool normalize)
{
	qtrk::guard("ComputeRadialProfile", [&] {
		std::vector<float> dirs(2 * (size_t)std::max(angularSteps, 1));
		for (int j = 0; j < angularSteps; j++) {
			float ang = 2 * 3.141593f * j / (float)angularSteps;
			dirs[2 * j] = cosf(ang);
			dirs[2 * j + 1] = sinf(ang);
		}
		const int w = src->w, h = src->h;
		const float* img = src->data;
		float rstep = (maxradius - minradius) / radialSteps;
		for (int i = 0; i < radialSteps; i++) {
			double sum = 0.0;
			int n = 0;
			float r = minradius + rstep * i;
			for (int a = 0; a < angularSteps; a++) {
				float x = center.x + dirs[2 * a] * r;
				float y = center.y + dirs[2 * a + 1] * r;
				if (w < 2 || h < 2 || !(x > -1.0f && y > -1.0f && x < (float)(w - 1) && y < (float)(h - 1))) continue;
				int rx = (int)x, ry = (int)y;
				const float* p = img + ry * w + rx;
				float fx = x - rx;
				float v0 = p[0] + (p[1] - p[0]) * fx;
				float v1 = p[w] + (p[w + 1] - p[w]) * fx;
				sum += v0 + (v1 - v0) * (y - ry);
				n++;
			}
			dst[i] = n > MIN_RADPROFILE_SMP_COUNT ? (float)(sum / n) : mean;
		}
		if (normalize) NormalizeRadialProfile(dst, radialSteps);
	});
}

// cputrack/utils.cpp:230-261
QTRK_API void NormalizeRadialProfile(float* prof, int rsteps)
{
	double sum = 0.0;
	for (int i = 0; i < rsteps; i++) sum += prof[i];
	float mean = (float)(sum / rsteps);
	double rms2 = 0.0;
	for (int i = 0; i < rsteps; i++) {
		prof[i] -= mean;
		rms2 += (float)(prof[i] * prof[i]);
	}
	double inv = 1.0f / sqrt(rms2 / rsteps);
	for (int i = 0; i < rsteps; i++) prof[i] = (float)(prof[i] * inv);
}

// cputrack/utils.h:23-34 (min-max scaling to [0,1])
QTRK_API void NormalizeImage(ImageData* image)
{
	const int n = image->w * image->h;
	if (n <= 0) return;
	float mx = image->data[0], mn = image->data[0];
	for (int k = 0; k < n; k++) {
		mx = std::max(mx, image->data[k]);
		mn = std::min(mn, image->data[k]);
	}
	for (int k = 0; k < n; k++) image->data[k] = (image->data[k] - mn) / (mx - mn);
}

// cputrack/CubicBSpline.h:15-24
static void bspline_weights(float w[4], float t)
{
	float t2 = t * t;
	float t3 = t2 * t;
	float omt = 1 - t;
	w[0] = 1.0f / 6 * omt * omt * omt;
	w[1] = 1.0f / 6 * (3 * t3 - 6 * t2 + 4);
	w[2] = 1.0f / 6 * (-3 * t3 + 3 * t2 + 3 * t + 1);
	w[3] = 1.0f / 6 * t3;
}

// Image of a bead from a radial LUT [planes = zlut->h][steps = zlut->w],
// cputrack/utils.cpp:325-390: B-spline (or linear) interpolation in z, linear
// in r, ovs x ovs sub-pixel samples with pixel centres at integer coordinates.
QTRK_API void GenerateImageFromLUT(ImageData* image, ImageData* zlut, float minradius, float maxradius, vector3f pos,
                                   bool useSplineInterp, int ovs)
{
	qtrk::guard("GenerateImageFromLUT", [&] {
		const int steps = zlut->w, planes = zlut->h;
		float radialcov = steps / (maxradius - minradius);
		std::vector<float> zi(steps);
		if (useSplineInterp) {
			int iz = std::max(1, std::min(planes - 3, (int)pos.z));
			float wt[4];
			bspline_weights(wt, pos.z - iz);
			for (int r = 0; r < steps; r++) {
				float v = 0;
				for (int i = 0; i < 4; i++) v += wt[i] * zlut->data[(size_t)(i - 1 + iz) * steps + r];
				zi[r] = v;
			}
		} else {
			int iz = (int)pos.z;
			for (int r = 0; r < steps; r++) {
				if (iz < 0) zi[r] = zlut->data[r];
				else if (iz >= planes - 1) zi[r] = zlut->data[(size_t)(planes - 1) * steps + r];
				else {
					float a = zlut->data[(size_t)iz * steps + r], b = zlut->data[(size_t)(iz + 1) * steps + r];
					zi[r] = a + (b - a) * (pos.z - iz);
				}
			}
		}
		float st = 1.0f / ovs;
		for (int y = 0; y < image->h; y++)
			for (int x = 0; x < image->w; x++) {
				float s = 0.0f;
				for (int ox = 0; ox < ovs; ox++)
					for (int oy = 0; oy < ovs; oy++) {
						float X = x + (ox + 0.5f) * st - pos.x - 0.5f;
						float Y = y + (oy + 0.5f) * st - pos.y - 0.5f;
						float pr = sqrtf(X * X + Y * Y);
						float r = (pr - minradius) * radialcov;
						if (r > steps - 2) r = (float)(steps - 2);
						if (r < 0) r = 0;
						int i = (int)r;
						s += zi[i] + (zi[i + 1] - zi[i]) * (r - i);
					}
				image->data[(size_t)y * image->w + x] = s / (ovs * ovs);
			}
	});
}

// Poisson noise on `poissonMax * pixel` expected counts, rescaled to maxValue
// (cputrack/utils.cpp:404-424).  Exact Poisson sampling (std::poisson_distribution)
// replaces the reference's rand()-based normal approximation above lambda=10.
QTRK_API void ApplyPoissonNoise(ImageData* img, float poissonMax, float maxValue)
{
	float ratio = maxValue / poissonMax;
	for (int i = 0; i < img->w * img->h; i++) {
		double lam = (double)(poissonMax * img->data[i]);
		int k = 0;
		if (lam > 0) {
			std::poisson_distribution<int> d(lam);
			k = d(rng());
		}
		img->data[i] = (float)(int)(k * ratio);
	}
}

// cputrack/utils.cpp:426-433
QTRK_API void ApplyGaussianNoise(ImageData* img, float sigma)
{
	std::normal_distribution<float> nd(0.0f, 1.0f);
	for (int k = 0; k < img->w * img->h; k++) {
		float v = img->data[k] + sigma * nd(rng());
		img->data[k] = v < 0.0f ? 0.0f : v;
	}
}

// cputrack/utils.cpp:393-402
QTRK_API void GenerateGaussianSpotImage(ImageData* img, vector2f pos, float sigma, float I0, float Ibg)
{
	float ed = 1.0f / sqrtf(2 * sigma * sigma);
	for (int y = 0; y < img->h; y++)
		for (int x = 0; x < img->w; x++) {
			float dx = 0.5f * erff((x - pos.x + .5f) * ed) - 0.5f * erff((x - pos.x - .5f) * ed);
			float dy = 0.5f * erff((y - pos.y + .5f) * ed) - 0.5f * erff((y - pos.y - .5f) * ed);
			img->data[(size_t)y * img->w + x] = Ibg + I0 * dx * dy;
		}
}

QTRK_API double GetPreciseTime(void)
{
	using namespace std::chrono;
	return duration<double>(steady_clock::now().time_since_epoch()).count();
}

QTRK_API void QTrkGenerateImagesFromLUT(float* out, int n, int width, int height, const float* luts, int nluts,
                                        int planes, int steps, float minradius, float maxradius, const float* pos,
                                        const int* lutIndex, bool useSplineInterp, int ovs)
{
	qtrk::guard("QTrkGenerateImagesFromLUT", [&] {
		for (int i = 0; i < n; i++)
			if (lutIndex[i] < 0 || lutIndex[i] >= nluts) qtrk::fail("lutIndex out of range");
		int nt = (int)std::thread::hardware_concurrency();
		nt = std::max(1, std::min(nt, 64));
		std::vector<std::thread> th;
		for (int t = 0; t < nt; t++)
			th.emplace_back([=] {
				for (int i = t; i < n; i += nt) {
					ImageData img{ out + (size_t)i * width * height, width, height };
					ImageData lut{ (float*)luts + (size_t)lutIndex[i] * planes * steps, steps, planes };
					vector3f p{ pos[3 * i], pos[3 * i + 1], pos[3 * i + 2] };
					GenerateImageFromLUT(&img, &lut, minradius, maxradius, p, useSplineInterp, ovs);
				}
			});
		for (auto& x : th) x.join();
	});
}
