// copy_pool.hpp -- host memcpy split over a few persistent threads.
//
// ScheduleFrame copies the caller's frame rows into pinned memory before it
// returns (the reference's ownership rule, cputrack/QueuedTracker.cpp:91-111);
// one thread copies at about 10 GB/s, far below what the host's memory and the
// PCIe link move, so large copies are split into parts that the caller and the
// pool's threads copy together.
#pragma once

#include <emmintrin.h>

#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace qtrk {

// memcpy with non-temporal stores: the destination lines are written without
// being read first, and are not kept in the caches (the GPU's DMA engine, not
// this CPU, reads them next).
inline void copy_streaming(void* dst, const void* src, size_t n)
{
	uint8_t* d = (uint8_t*)dst;
	const uint8_t* s = (const uint8_t*)src;
	const size_t head = (16 - ((uintptr_t)d & 15)) & 15;
	if (n <= head + 64) {
		std::memcpy(d, s, n);
		return;
	}
	std::memcpy(d, s, head);
	d += head;
	s += head;
	n -= head;
	for (size_t i = 0, m = n / 64; i < m; i++, d += 64, s += 64) {
		const __m128i a = _mm_loadu_si128((const __m128i*)s), b = _mm_loadu_si128((const __m128i*)(s + 16));
		const __m128i c = _mm_loadu_si128((const __m128i*)(s + 32)), e = _mm_loadu_si128((const __m128i*)(s + 48));
		_mm_stream_si128((__m128i*)d, a);
		_mm_stream_si128((__m128i*)(d + 16), b);
		_mm_stream_si128((__m128i*)(d + 32), c);
		_mm_stream_si128((__m128i*)(d + 48), e);
	}
	_mm_sfence();
	std::memcpy(d, s, n & 63);
}

class CopyPool {
public:
	explicit CopyPool(int helpers, bool streaming = false) : helpers_(helpers < 0 ? 0 : helpers), streaming_(streaming)
	{
		try {
			for (int i = 0; i < helpers_; i++) threads_.emplace_back([this] { Run(); });
		} catch (...) {
			Stop();  // the threads already started must end before the members go
			throw;
		}
	}
	~CopyPool() { Stop(); }
	CopyPool(const CopyPool&) = delete;
	CopyPool& operator=(const CopyPool&) = delete;

	// Copies `bytes` from src to dst and returns when all of it is done.  Calls
	// from several threads take turns.
	void Copy(void* dst, const void* src, size_t bytes)
	{
		if (bytes < kMinSplit || helpers_ == 0) {
			if (streaming_)
				copy_streaming(dst, src, bytes);
			else
				std::memcpy(dst, src, bytes);
			return;
		}
		std::lock_guard<std::mutex> serial(call_);
		const int parts = helpers_ + 1;
		// 4 KB-aligned parts, so that neighbouring parts do not share pages
		const size_t chunk = ((bytes + parts - 1) / parts + 4095) & ~(size_t)4095;
		{
			std::lock_guard<std::mutex> g(m_);
			dst_ = (uint8_t*)dst;
			src_ = (const uint8_t*)src;
			bytes_ = bytes;
			chunk_ = chunk;
			next_ = 1;  // part 0 is the caller's
			parts_ = parts;
			done_ = 0;
			gen_++;
		}
		cv_.notify_all();
		CopyPart(0);
		std::unique_lock<std::mutex> lk(m_);
		done_++;
		done_cv_.wait(lk, [&] { return done_ == parts_; });
	}

private:
	static constexpr size_t kMinSplit = 1 << 20;

	void Stop()
	{
		{
			std::lock_guard<std::mutex> g(m_);
			stop_ = true;
		}
		cv_.notify_all();
		for (auto& t : threads_) t.join();
		threads_.clear();
	}

	void CopyPart(int p)
	{
		const size_t off = (size_t)p * chunk_;
		if (off >= bytes_) return;
		const size_t n = off + chunk_ > bytes_ ? bytes_ - off : chunk_;
		if (streaming_) copy_streaming(dst_ + off, src_ + off, n);
		else std::memcpy(dst_ + off, src_ + off, n);
	}

	void Run()
	{
		uint64_t seen = 0;
		std::unique_lock<std::mutex> lk(m_);
		while (true) {
			cv_.wait(lk, [&] { return stop_ || (gen_ != seen && next_ < parts_); });
			if (stop_) return;
			seen = gen_;
			while (next_ < parts_) {
				const int p = next_++;
				lk.unlock();
				CopyPart(p);
				lk.lock();
				if (++done_ == parts_) done_cv_.notify_all();
			}
		}
	}

	int helpers_;
	bool streaming_;
	std::vector<std::thread> threads_;
	std::mutex call_, m_;
	std::condition_variable cv_, done_cv_;
	uint8_t* dst_ = nullptr;
	const uint8_t* src_ = nullptr;
	size_t bytes_ = 0, chunk_ = 0;
	int next_ = 0, parts_ = 0, done_ = 0;
	uint64_t gen_ = 0;
	bool stop_ = false;
};

}  // namespace qtrk
