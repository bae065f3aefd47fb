/*
 * ResultManager.h -- the result sink after FetchResults (cputrack/ResultManager.h:10-69).
 *
 * A background thread drains a tracker's results into per-frame tables (one
 * slot per bead), keeps frame counters, and appends finished frames to a text
 * or binary (version 3) trace file every writeInterval frames.  The file
 * formats are the reference's byte for byte; see result_manager.cpp.
 */
#ifndef QTRK_RESULT_MANAGER_H
#define QTRK_RESULT_MANAGER_H

#include "QueuedTracker.h"

#include <atomic>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

class ResultManager {
public:
	ResultManager(const char* outfile, const char* frameinfo, ResultManagerConfig* cfg,
	              std::vector<std::string> colnames);
	~ResultManager();

	void SetTracker(QueuedTracker* qtrk);
	QueuedTracker* GetTracker();

	int GetBeadPositions(int startFrame, int endFrame, int bead, LocalizationResult* r);
	int GetResults(LocalizationResult* results, int startFrame, int numFrames);
	void Flush();

	RMFrameCounters GetFrameCounters();
	void StoreFrameInfo(int frame, double timestamp, float* columns);
	int GetFrameCount();

	bool RemoveBeadResults(int bead);
	const ResultManagerConfig& Config() { return config; }

protected:
	struct FrameResult {
		FrameResult(int nResult, int nFrameInfo) : results(nResult), frameInfo(nFrameInfo) {}
		std::vector<LocalizationResult> results;
		std::vector<float> frameInfo;
		int count = 0;
		double timestamp = 0;
		bool hasFrameInfo = false;
	};

	bool CheckResultSpace(int fr);
	void Write();
	void WriteBinaryResults();
	void WriteTextResults();
	void WriteBinaryFileHeader();
	void StoreResult(const LocalizationResult* r);
	bool Update();
	void ThreadLoop();

	std::recursive_mutex resultMutex;
	std::mutex trackerMutex;
	std::vector<std::string> frameInfoNames;
	std::deque<std::unique_ptr<FrameResult>> frameResults;
	RMFrameCounters cnt{};
	ResultManagerConfig config;
	QueuedTracker* qtrk = nullptr;
	std::vector<LocalizationResult> fetchBuf;  // used under trackerMutex
	std::string outputFile, frameInfoFile;
	std::atomic<bool> quit{ false };
	std::thread thread;
};

#endif
