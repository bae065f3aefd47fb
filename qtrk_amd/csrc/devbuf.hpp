// devbuf.hpp -- owning GPU and pinned-host buffers that grow on demand.
#pragma once

#include <cstddef>

namespace qtrk {

// GPU memory (hipMalloc); ensure(n) reallocates only when n exceeds the size held.
struct DeviceBuffer {
	void* p = nullptr;
	size_t bytes = 0;
	DeviceBuffer() = default;
	DeviceBuffer(const DeviceBuffer&) = delete;
	DeviceBuffer& operator=(const DeviceBuffer&) = delete;
	void ensure(size_t n);
	void release();
	template <typename T> T* as() const { return (T*)p; }
	~DeviceBuffer() { release(); }
};

// Page-locked host memory (hipHostMalloc), for asynchronous copies.
struct PinnedBuffer {
	void* p = nullptr;
	size_t bytes = 0;
	PinnedBuffer() = default;
	PinnedBuffer(const PinnedBuffer&) = delete;
	PinnedBuffer& operator=(const PinnedBuffer&) = delete;
	void ensure(size_t n);
	void release();
	template <typename T> T* as() const { return (T*)p; }
	~PinnedBuffer() { release(); }
};

}  // namespace qtrk
