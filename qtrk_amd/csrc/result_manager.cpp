// result_manager.cpp -- the result sink after FetchResults and its RM* C exports
// (cputrack/ResultManager.cpp, cputrack/qtrk_c_api.cpp:181-241).
//
// Per-frame tables hold one result per bead, with positions relative to the ROI
// centre, (pos - (w/2, h/2, 0) + offset) * scaling.  A frame is processed when
// every bead has reported, or given up (counted lost) once more than 1,000
// frames of frame info have arrived after it.  Every writeInterval processed
// frames the new ones are appended to the output file:
//
//   text:   "<frame>\t<timestamp %f>\t" then "%.7f\t%.7f\t%.7f\t" per bead, "\n";
//           the frame-info file gets "<frame>\t<timestamp>\t" + "%.7f\t" per column.
//   binary: header int32 version (3), int32 numBeads, int32 numFrameInfoColumns,
//           int32 byte offset of the frame data, the column names NUL-terminated;
//           per frame uint32 frame, float64 timestamp, float32 info[columns],
//           float32 pos[beads][3], int32 error[beads], float32 imageMean[beads].
//
// The data offset is 4 bytes: the reference's `long` on its Windows (LLP64)
// build, which is what its trace reader takes (TraceViewer/ViewTraceDlg.cs:208-230).
// Where the reference would index outside its tables (bead indices beyond
// numBeads, frames dropped from memory before they were written, a removed
// bead), these tables are bounds-checked: such results are counted as lost and
// such frames skipped.
#include "ResultManager.h"
#include "errors.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>

using qtrk::fail;

namespace {
constexpr int kVersion = 3;
constexpr int kFetch = 1024;  // results taken from the tracker per update
constexpr int kIdleMs = 5;    // poll period while the tracker has nothing
}  // namespace

ResultManager::ResultManager(const char* outfile, const char* frameinfo, ResultManagerConfig* cfg,
                             std::vector<std::string> colnames)
{
	config = *cfg;
	outputFile = outfile ? outfile : "";
	frameInfoFile = frameinfo ? frameinfo : "";
	frameInfoNames = std::move(colnames);
	if (!outputFile.empty()) remove(outputFile.c_str());
	if (!frameInfoFile.empty()) remove(frameInfoFile.c_str());
	if (config.binaryOutput) WriteBinaryFileHeader();
	thread = std::thread([this] { ThreadLoop(); });
}

ResultManager::~ResultManager()
{
	quit = true;
	if (thread.joinable()) thread.join();
}

void ResultManager::WriteBinaryFileHeader()
{
	FILE* f = fopen(outputFile.c_str(), "wb");
	if (!f) fail("Unable to open file " + outputFile);
	const int head[4] = { kVersion, config.numBeads, config.numFrameInfoColumns, 1234 };
	fwrite(head, sizeof(int), 4, f);
	for (int i = 0; i < config.numFrameInfoColumns; i++) {
		const std::string n = i < (int)frameInfoNames.size() ? frameInfoNames[i] : std::string();
		fwrite(n.c_str(), n.length() + 1, 1, f);
	}
	const int32_t offset = (int32_t)ftell(f);
	fseek(f, 3 * sizeof(int), SEEK_SET);
	fwrite(&offset, sizeof(offset), 1, f);
	fclose(f);
}

void ResultManager::SetTracker(QueuedTracker* t)
{
	std::lock_guard<std::mutex> g(trackerMutex);
	qtrk = t;
}

QueuedTracker* ResultManager::GetTracker()
{
	std::lock_guard<std::mutex> g(trackerMutex);
	return qtrk;
}

bool ResultManager::CheckResultSpace(int fr)
{
	if (fr < cnt.startFrame) return false;                 // already dropped from memory
	if (fr > cnt.processedFrames + 20000) return false;    // a suspiciously large frame number
	while (fr >= cnt.startFrame + (int)frameResults.size())
		frameResults.emplace_back(new FrameResult(config.numBeads, config.numFrameInfoColumns));
	return true;
}

// Caller holds trackerMutex (qtrk is valid) and resultMutex.
void ResultManager::StoreResult(const LocalizationResult* r)
{
	const int frame = (int)r->job.frame;
	if (!CheckResultSpace(frame)) {
		cnt.lostFrames++;
		return;
	}
	FrameResult* fr = frameResults[frame - cnt.startFrame].get();
	const int bead = r->job.zlutIndex;
	if (bead < 0 || bead >= (int)fr->results.size()) {
		cnt.lostFrames++;
		return;
	}
	LocalizationResult s = *r;
	s.pos.x = ((s.pos.x - qtrk->cfg.width * 0.5f) + config.offset.x) * config.scaling.x;
	s.pos.y = ((s.pos.y - qtrk->cfg.height * 0.5f) + config.offset.y) * config.scaling.y;
	s.pos.z = ((s.pos.z - 0.0f) + config.offset.z) * config.scaling.z;
	fr->results[bead] = s;
	fr->count++;
	// frames dropped from memory before they completed count as lost
	if (cnt.processedFrames < cnt.startFrame) {
		cnt.lostFrames += cnt.startFrame - cnt.processedFrames;
		cnt.processedFrames = cnt.startFrame;
	}
	while (cnt.processedFrames - cnt.startFrame < (int)frameResults.size()) {
		const FrameResult* p = frameResults[cnt.processedFrames - cnt.startFrame].get();
		if (!(p->count == config.numBeads || cnt.capturedFrames - cnt.processedFrames > 1000)) break;
		if (p->count < config.numBeads) cnt.lostFrames++;
		cnt.processedFrames++;
	}
	cnt.localizationsDone++;
}

void ResultManager::WriteBinaryResults()
{
	if (outputFile.empty()) return;
	FILE* f = fopen(outputFile.c_str(), "ab");
	if (!f) {
		cnt.fileError++;
		return;
	}
	for (int j = std::max(cnt.lastSaveFrame, cnt.startFrame); j < cnt.processedFrames; j++) {
		const FrameResult* fr = frameResults[j - cnt.startFrame].get();
		const uint32_t id = (uint32_t)j;
		fwrite(&id, sizeof(id), 1, f);
		fwrite(&fr->timestamp, sizeof(double), 1, f);
		if (config.numFrameInfoColumns > 0) fwrite(fr->frameInfo.data(), sizeof(float), config.numFrameInfoColumns, f);
		const int nb = (int)fr->results.size();
		for (int i = 0; i < config.numBeads; i++) {
			const vector3f p = i < nb ? fr->results[i].pos : vector3f{ 0, 0, 0 };
			fwrite(&p, sizeof(vector3f), 1, f);
		}
		for (int i = 0; i < config.numBeads; i++) {
			const uint32_t e = i < nb ? fr->results[i].error : 0;
			fwrite(&e, sizeof(e), 1, f);
		}
		for (int i = 0; i < config.numBeads; i++) {
			const float m = i < nb ? fr->results[i].imageMean : 0.0f;
			fwrite(&m, sizeof(m), 1, f);
		}
	}
	fclose(f);
}

void ResultManager::WriteTextResults()
{
	FILE* f = outputFile.empty() ? nullptr : fopen(outputFile.c_str(), "a");
	FILE* finfo = frameInfoFile.empty() ? nullptr : fopen(frameInfoFile.c_str(), "a");
	if ((!outputFile.empty() && !f) || (!frameInfoFile.empty() && !finfo)) cnt.fileError++;
	for (int k = std::max(cnt.lastSaveFrame, cnt.startFrame); k < cnt.processedFrames; k++) {
		const FrameResult* fr = frameResults[k - cnt.startFrame].get();
		if (f) {
			fprintf(f, "%d\t%f\t", k, fr->timestamp);
			const int nb = (int)fr->results.size();
			for (int i = 0; i < config.numBeads; i++) {
				const vector3f p = i < nb ? fr->results[i].pos : vector3f{ 0, 0, 0 };
				fprintf(f, "%.7f\t%.7f\t%.7f\t", p.x, p.y, p.z);
			}
			fputs("\n", f);
		}
		if (finfo) {
			fprintf(finfo, "%d\t%f\t", k, fr->timestamp);
			for (int i = 0; i < config.numFrameInfoColumns; i++) fprintf(finfo, "%.7f\t", fr->frameInfo[i]);
			fputs("\n", finfo);
		}
	}
	if (finfo) fclose(finfo);
	if (f) fclose(f);
}

void ResultManager::Write()
{
	std::lock_guard<std::recursive_mutex> g(resultMutex);
	if (config.binaryOutput) WriteBinaryResults();
	else WriteTextResults();
	cnt.lastSaveFrame = cnt.processedFrames;
}

// One drain of the tracker (cputrack/ResultManager.cpp:227-278); true if any result came.
bool ResultManager::Update()
{
	std::unique_lock<std::mutex> tl(trackerMutex);
	if (!qtrk) return false;
	fetchBuf.resize(kFetch);
	const int count = qtrk->FetchResults(fetchBuf.data(), kFetch);
	std::lock_guard<std::recursive_mutex> rl(resultMutex);
	for (int i = 0; i < count; i++) StoreResult(&fetchBuf[i]);
	tl.unlock();

	if (cnt.processedFrames - cnt.lastSaveFrame >= config.writeInterval) Write();

	if (config.maxFramesInMemory > 0 && frameResults.size() > config.maxFramesInMemory) {
		const int del = (int)(frameResults.size() - config.maxFramesInMemory);
		if (cnt.processedFrames < cnt.startFrame) {
			// unfinished frames are written out with the results they have
			const int lost = cnt.startFrame - cnt.processedFrames;
			cnt.processedFrames += lost;
			cnt.lostFrames += lost;
			Write();
		}
		frameResults.erase(frameResults.begin(), frameResults.begin() + del);
		cnt.startFrame += del;
	}
	return count > 0;
}

void ResultManager::ThreadLoop()
{
	while (true) {
		if (!Update()) std::this_thread::sleep_for(std::chrono::milliseconds(kIdleMs));
		if (quit) break;
	}
}

int ResultManager::GetBeadPositions(int startfr, int end, int bead, LocalizationResult* results)
{
	int count = end - startfr;
	std::lock_guard<std::recursive_mutex> g(resultMutex);
	int start = startfr - cnt.startFrame;
	if (start < 0) start = 0;
	if (count > cnt.processedFrames - cnt.startFrame) count = cnt.processedFrames - cnt.startFrame;
	count = std::min(count, (int)frameResults.size() - start);
	if (count < 0) count = 0;
	for (int i = 0; i < count; i++) {
		const FrameResult* fr = frameResults[i + start].get();
		results[i] = bead >= 0 && bead < (int)fr->results.size() ? fr->results[bead] : LocalizationResult{};
	}
	return count;
}

// Drains what the tracker has finished after flushing it, then writes every processed frame.
void ResultManager::Flush()
{
	{
		std::lock_guard<std::mutex> g(trackerMutex);
		if (qtrk) qtrk->Flush();
	}
	while (Update()) {
	}
	Write();
}

RMFrameCounters ResultManager::GetFrameCounters()
{
	std::lock_guard<std::recursive_mutex> g(resultMutex);
	return cnt;
}

int ResultManager::GetResults(LocalizationResult* results, int startFrame, int numFrames)
{
	std::lock_guard<std::recursive_mutex> g(resultMutex);
	if (!(startFrame >= cnt.startFrame && numFrames + startFrame <= cnt.processedFrames)) return 0;
	for (int f = 0; f < numFrames; f++) {
		const FrameResult* fr = frameResults[f + startFrame - cnt.startFrame].get();
		for (int j = 0; j < config.numBeads; j++)
			results[(size_t)config.numBeads * f + j] = j < (int)fr->results.size() ? fr->results[j] : LocalizationResult{};
	}
	return numFrames;
}

void ResultManager::StoreFrameInfo(int frame, double timestamp, float* columns)
{
	std::lock_guard<std::recursive_mutex> g(resultMutex);
	if (!CheckResultSpace(frame)) return;
	FrameResult* fr = frameResults[frame - cnt.startFrame].get();
	if (fr->hasFrameInfo) return;
	fr->timestamp = timestamp;
	for (int i = 0; i < config.numFrameInfoColumns; i++) fr->frameInfo[i] = columns ? columns[i] : 0.0f;
	++cnt.capturedFrames;
	fr->hasFrameInfo = true;
}

int ResultManager::GetFrameCount()
{
	std::lock_guard<std::recursive_mutex> g(resultMutex);
	return cnt.capturedFrames;
}

// Drops one bead's slot from every frame in memory, as the reference does; the file
// already written keeps it and config.numBeads is unchanged (cputrack/ResultManager.cpp:386-399).
bool ResultManager::RemoveBeadResults(int bead)
{
	std::lock_guard<std::recursive_mutex> g(resultMutex);
	for (auto& fr : frameResults) {
		fr->count--;
		if (bead >= 0 && bead < (int)fr->results.size()) fr->results.erase(fr->results.begin() + bead);
	}
	return true;
}

// ---------------------------------------------------------------- C exports

QTRK_API ResultManager* RMCreate(const char* file, const char* frameinfo, ResultManagerConfig* cfg,
                                 const char* semicolonSeparatedNames)
{
	ResultManager* rm = nullptr;
	GUARD({
		if (!cfg) fail("null config");
		if (cfg->numBeads < 0 || cfg->numFrameInfoColumns < 0) fail("negative numBeads or numFrameInfoColumns");
		// The reference's split (cputrack/qtrk_c_api.cpp:183-193) appends the ';' itself to the
		// next name ("a;b" -> "a", ";b"); the names go into the binary header as they are.
		std::vector<std::string> colnames;
		std::string name;
		for (int i = 0; semicolonSeparatedNames && semicolonSeparatedNames[i]; i++) {
			const char c = semicolonSeparatedNames[i];
			if (c == ';') {
				colnames.push_back(name);
				name.clear();
			}
			name += c;
		}
		if (!name.empty()) colnames.push_back(name);
		rm = new ResultManager(file, frameinfo, cfg, colnames);
	});
	return rm;
}

static ResultManager* checked(ResultManager* rm)
{
	if (!rm) fail("null result manager");
	return rm;
}

QTRK_API void RMSetTracker(ResultManager* rm, QueuedTracker* qtrk)
{
	GUARD({ checked(rm)->SetTracker(qtrk); });
}

QTRK_API void RMDestroy(ResultManager* rm)
{
	GUARD({ delete rm; });
}

QTRK_API void RMStoreFrameInfo(ResultManager* rm, int frame, double timestamp, float* cols)
{
	GUARD({ checked(rm)->StoreFrameInfo(frame, timestamp, cols); });
}

QTRK_API int RMGetBeadResults(ResultManager* rm, int start, int numFrames, int bead, LocalizationResult* results)
{
	int n = 0;
	GUARD({
		if (numFrames > 0 && !results) fail("null results");
		n = checked(rm)->GetBeadPositions(start, start + numFrames, bead, results);
	});
	return n;
}

QTRK_API void RMGetFrameCounters(ResultManager* rm, RMFrameCounters* fc)
{
	GUARD({
		if (!fc) fail("null counters");
		*fc = checked(rm)->GetFrameCounters();
	});
}

QTRK_API void RMFlush(ResultManager* rm)
{
	GUARD({ checked(rm)->Flush(); });
}

QTRK_API int RMGetResults(ResultManager* rm, int startFrame, int numFrames, LocalizationResult* results)
{
	int n = 0;
	GUARD({
		if (numFrames > 0 && !results) fail("null results");
		n = checked(rm)->GetResults(results, startFrame, numFrames);
	});
	return n;
}

QTRK_API void RMRemoveBead(ResultManager* rm, int bead)
{
	GUARD({ checked(rm)->RemoveBeadResults(bead); });
}

QTRK_API void RMGetConfig(ResultManager* rm, ResultManagerConfig* cfg)
{
	GUARD({
		if (!cfg) fail("null config");
		*cfg = checked(rm)->Config();
	});
}
