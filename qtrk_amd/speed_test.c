/*
 * qtrk_speed_test -- the reference's speed-test clock through the C ABI, as a C client (the role
 * QTrkDotNet or the LabVIEW layer plays) sees it: from the first schedule call until
 * QTrkGetResultCount() reaches N after QTrkFlush (cudatrack-test/test.cu:310-371 SpeedTest,
 * cputrack-test/SharedTests.h:300-314 WaitForFinish).  The ROIs come from host memory one call
 * each, in one of two ways:
 *   frame1  QTrkScheduleFrame with the ROI as a one-ROI frame at (0, 0) -- SpeedTest's own call
 *   roi     QTrkScheduleLocalization per ROI
 * The wait polls the result count every 20 us (WaitForFinish sleeps 50 ms between polls, which
 * would be most of the time here).
 *
 *   qtrk_speed_test <dir> <frame1|roi> <passes>
 *
 * <dir> holds params.txt ("n width height luts planes steps locmode"), settings.bin (QTrkSettings),
 * lut.bin (float [luts][planes][steps]), rois.bin (uint16 [n][height][width]) and jobs.bin
 * (LocalizationJob [n]).  Writes results.bin (LocalizationResult [n], the last pass, in fetch
 * order) and prints one JSON line: per pass, loc/s and the mean time of a schedule call.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "qtrk_c_api.h"

static double now_s(void)
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void* read_file(const char* dir, const char* name, size_t bytes)
{
	char path[4096];
	snprintf(path, sizeof path, "%s/%s", dir, name);
	FILE* f = fopen(path, "rb");
	if (!f) {
		fprintf(stderr, "cannot open %s\n", path);
		exit(2);
	}
	void* p = malloc(bytes ? bytes : 1);
	if (!p || fread(p, 1, bytes, f) != bytes) {
		fprintf(stderr, "%s: expected %zu bytes\n", path, bytes);
		exit(2);
	}
	fclose(f);
	return p;
}

static void sleep_us(long us)
{
	struct timespec t = { 0, us * 1000L };
	nanosleep(&t, NULL);
}

int main(int argc, char** argv)
{
	if (argc != 4 || (strcmp(argv[2], "frame1") != 0 && strcmp(argv[2], "roi") != 0)) {
		fprintf(stderr, "usage: %s <dir> <frame1|roi> <passes>\n", argv[0]);
		return 2;
	}
	const char* dir = argv[1];
	const int frame1 = strcmp(argv[2], "frame1") == 0;
	const int passes = atoi(argv[3]);
	int n, w, h, luts, planes, steps, mode;
	{
		char path[4096];
		snprintf(path, sizeof path, "%s/params.txt", dir);
		FILE* f = fopen(path, "r");
		if (!f || fscanf(f, "%d %d %d %d %d %d %d", &n, &w, &h, &luts, &planes, &steps, &mode) != 7) {
			fprintf(stderr, "cannot read %s\n", path);
			return 2;
		}
		fclose(f);
	}
	if (n <= 0 || w <= 0 || h <= 0 || passes <= 0) {
		fprintf(stderr, "invalid parameters\n");
		return 2;
	}
	QTrkSettings* cfg = (QTrkSettings*)read_file(dir, "settings.bin", sizeof(QTrkSettings));
	float* lut = (float*)read_file(dir, "lut.bin", sizeof(float) * (size_t)luts * planes * steps);
	uint16_t* rois = (uint16_t*)read_file(dir, "rois.bin", sizeof(uint16_t) * (size_t)n * w * h);
	LocalizationJob* jobs = (LocalizationJob*)read_file(dir, "jobs.bin", sizeof(LocalizationJob) * (size_t)n);
	LocalizationResult* res = (LocalizationResult*)malloc(sizeof(LocalizationResult) * (size_t)n);

	QueuedTracker* trk = QTrkCreateInstance(cfg);
	if (!trk) {
		fprintf(stderr, "QTrkCreateInstance failed\n");
		return 1;
	}
	QTrkSetRadialZLUT(trk, lut, luts, planes);
	QTrkSetLocalizationMode(trk, (LocMode_t)mode);
	ROIPosition at0 = { 0, 0 };
	const size_t roi_px = (size_t)w * h;

	printf("{\"mode\": \"%s\", \"rois\": %d, \"locs_per_s\": [", argv[2], n);
	double* sched = (double*)malloc(sizeof(double) * (size_t)passes);
	int got = 0;
	for (int p = 0; p < passes; p++) {
		QTrkClearResults(trk);
		const double t0 = now_s();
		for (int i = 0; i < n; i++) {
			if (frame1)
				QTrkScheduleFrame(trk, rois + i * roi_px, w * 2, w, h, &at0, 1, QTrkU16, &jobs[i]);
			else
				QTrkScheduleLocalization(trk, rois + i * roi_px, w * 2, QTrkU16, &jobs[i]);
		}
		const double t1 = now_s();
		QTrkFlush(trk);
		while (QTrkGetResultCount(trk) < n) sleep_us(20);
		const double t2 = now_s();
		sched[p] = t1 - t0;
		got = QTrkFetchResults(trk, res, n);
		printf("%s%.1f", p ? ", " : "", n / (t2 - t0));
	}
	printf("], \"schedule_call_us\": [");
	for (int p = 0; p < passes; p++) printf("%s%.3f", p ? ", " : "", 1e6 * sched[p] / n);
	printf("], \"fetched\": %d}\n", got);
	free(sched);
	{
		char path[4096];
		snprintf(path, sizeof path, "%s/results.bin", dir);
		FILE* f = fopen(path, "wb");
		if (!f || fwrite(res, sizeof(LocalizationResult), (size_t)got, f) != (size_t)got) {
			fprintf(stderr, "cannot write %s\n", path);
			return 1;
		}
		fclose(f);
	}
	QTrkFreeInstance(trk);
	free(cfg);
	free(lut);
	free(rois);
	free(jobs);
	free(res);
	return 0;
}
