// cxx_surface_check.cpp -- the C++ members of QueuedTracker that the reference's LabVIEW
// wrapper calls (cputrack/lv_queuetrk_api.cpp:461, 478, 613; declared at
// cputrack/QueuedTracker.h:38, 87-92), on the GPU tracker made by CreateQueuedTracker:
//   * ComputeZBiasCorrection(int, CImageData*, int, bool) gives the table of the C export
//     QTrkComputeZLUTBiasTable bit for bit, and makes it the active table;
//   * GetZLUTBiasCorrection() returns a copy of the active table (NULL once SetRadialZLUT drops it);
//     SetZLUTBiasCorrection(const CImageData&) sets one, which QTrkGetZLUTBiasTable reads back;
//   * the 7-argument ScheduleLocalization and ScheduleImageData give the results of the
//     4-argument ScheduleLocalization on the same ROIs, with the job fields they were given;
//   * DebugImage returns an empty image.
// Prints "ok" and exits 0 when every check holds.  Built by the library's build step against
// include/ and qtrk_amd/libqtrk.so.
#include "ImageData.h"
#include "QueuedTracker.h"
#include "qtrk_amd.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

static int fails = 0;
#define CHECK(c)                                                        \
	do {                                                                \
		if (!(c)) {                                                     \
			std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);   \
			fails++;                                                    \
		}                                                               \
	} while (0)

int main()
{
	QTrkSettings s;
	QTrkGetDefaultConfig(&s);
	s.width = s.height = 64;
	s.qi_iterations = 3;
	const QTrkComputedConfig cc = QTrkComputeConfig(s);
	QueuedTracker* t = CreateQueuedTracker(cc);
	if (!t) {
		std::printf("FAIL: CreateQueuedTracker returned NULL: %s\n", QTrkGetLastError());
		return 1;
	}
	const int beads = 2, planes = 40, nb = t->cfg.zlut_radialsteps;
	const float minr = t->cfg.zlut_minradius, maxr = t->cfg.zlut_maxradius;
	std::vector<float> lut((size_t)beads * planes * nb);
	for (int b = 0; b < beads; b++)
		for (int p = 0; p < planes; p++)
			for (int i = 0; i < nb; i++) {
				const double r = minr + i * (maxr - minr) / nb, w = 5.0 + 0.08 * p, T = 7.0 + 0.05 * p + b;
				lut[((size_t)b * planes + p) * nb + i] = (float)(std::exp(-r * r / (2 * w * w)) * std::cos(2 * M_PI * r / T));
			}
	t->SetRadialZLUT(lut.data(), beads, planes);
	t->SetLocalizationMode(LT_QI | LT_LocalizeZ | LT_NormalizeProfile);

	// ComputeZBiasCorrection through the member and through the C export
	const int bp = 20;
	CImageData ci;
	t->ComputeZBiasCorrection(bp, &ci, 4, true);
	CHECK(ci.w == bp && ci.h == beads && ci.data != nullptr);
	std::vector<float> tab((size_t)bp * beads);
	CHECK(QTrkComputeZLUTBiasTable(t, bp, tab.data(), 4, true) == 0);
	CHECK(ci.data && std::memcmp(tab.data(), ci.data, tab.size() * 4) == 0);
	CImageData* g = t->GetZLUTBiasCorrection();
	CHECK(g && g->w == bp && g->h == beads && std::memcmp(g->data, tab.data(), tab.size() * 4) == 0);
	delete g;

	// SetZLUTBiasCorrection(const CImageData&) and both readers
	CImageData m(bp, beads);
	for (int i = 0; i < bp * beads; i++) m.data[i] = 0.01f * i - 0.2f;
	t->SetZLUTBiasCorrection(m);
	g = t->GetZLUTBiasCorrection();
	CHECK(g && g->w == bp && g->h == beads && std::memcmp(g->data, m.data, tab.size() * 4) == 0);
	delete g;
	int gp = 0, gc = 0;
	std::vector<float> back(tab.size());
	CHECK(QTrkGetZLUTBiasTable(t, back.data(), &gp, &gc) == 1 && gp == bp && gc == beads);
	CHECK(std::memcmp(back.data(), m.data, back.size() * 4) == 0);
	t->SetRadialZLUT(lut.data(), beads, planes);  // drops the table, as the reference does
	CHECK(t->GetZLUTBiasCorrection() == nullptr);

	// the 7-argument ScheduleLocalization and ScheduleImageData against the 4-argument form
	const int n = 6;
	std::vector<ImageData> rois;
	for (int i = 0; i < n; i++) {
		ImageData img = ImageData::alloc(64, 64);
		ImageData zl(&lut[(size_t)(i % beads) * planes * nb], nb, planes);
		vector3f pos = { 32.0f + 0.7f * i - 2.0f, 31.5f - 0.4f * i, 8.0f + 4.5f * i };
		GenerateImageFromLUT(&img, &zl, minr, maxr, pos, true, 4);
		for (int k = 0; k < 64 * 64; k++) img.data[k] = 100.0f + 1000.0f * img.data[k];
		rois.push_back(img);
	}
	for (int i = 0; i < n; i++) {
		LocalizationJob j;
		std::memset(&j, 0, sizeof(j));
		j.frame = i;
		j.timestamp = 1000 + i;
		j.zlutIndex = i % beads;
		if (i % 3 == 0) t->ScheduleLocalization(rois[i].data, rois[i].pitch(), QTrkFloat, &j);
		else if (i % 3 == 1) {
			vector3f init = { 1.0f, 2.0f, 3.0f };
			t->ScheduleLocalization((uchar*)rois[i].data, rois[i].pitch(), QTrkFloat, (uint)i, (uint)(1000 + i), &init,
			                        (uint)(i % beads));
		} else
			t->ScheduleImageData(&rois[i], &j);
		j.frame = 100 + i;
		t->ScheduleLocalization(rois[i].data, rois[i].pitch(), QTrkFloat, &j);
	}
	t->Flush();
	const auto t0 = std::chrono::steady_clock::now();
	while (t->GetResultCount() < 2 * n && std::chrono::steady_clock::now() - t0 < std::chrono::seconds(30))
		std::this_thread::sleep_for(std::chrono::milliseconds(1));
	std::vector<LocalizationResult> res(2 * n);
	CHECK(t->FetchResults(res.data(), 2 * n) == 2 * n);
	for (int i = 0; i < n; i++) {
		const LocalizationResult *a = nullptr, *b = nullptr;
		for (const LocalizationResult& r : res) {
			if ((int)r.job.frame == i) a = &r;
			if ((int)r.job.frame == 100 + i) b = &r;
		}
		CHECK(a && b);
		if (!a || !b) continue;
		CHECK(a->job.timestamp == (uint)(1000 + i) && a->job.zlutIndex == i % beads);
		CHECK(a->pos.x == b->pos.x && a->pos.y == b->pos.y && a->pos.z == b->pos.z);
		if (i % 3 == 1) CHECK(a->job.initialPos.x == 1.0f && a->job.initialPos.z == 3.0f);
	}
	for (ImageData& r : rois) r.free();

	ImageData d = t->DebugImage(0);
	CHECK(d.data == nullptr && d.w == 0 && d.h == 0);
	delete t;
	std::printf(fails ? "FAILED (%d)\n" : "ok\n", fails);
	return fails ? 1 : 0;
}
