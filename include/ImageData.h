/*
 * ImageData.h -- the image types of the C++ surface: TImageData<T> (a w x h
 * image of T in row-major order; does not own its data) and CImageData (a
 * float image that owns it), with the member names and meaning of the
 * reference's (cputrack/utils.h:77-161), which C++ callers such as the LabVIEW
 * wrappers use (cputrack/lv_queuetrk_api.cpp:461-478).  The layout -- data
 * pointer, width, height -- is the one the C ABI's ImageData has.
 */
#ifndef QTRK_IMAGE_DATA_H
#define QTRK_IMAGE_DATA_H

#include "qtrk_c_api.h"

#include <cstddef>

template <typename T>
struct TImageData {
	T* data;
	int w, h;

	TImageData() : data(nullptr), w(0), h(0) {}
	TImageData(T* d, int width, int height) : data(d), w(width), h(height) {}

	int numPixels() const { return w * h; }
	int pitch() const { return (int)sizeof(T) * w; } /* bytes per row */
	T& at(int x, int y) { return data[(size_t)w * y + x]; }
	T& operator[](int i) { return data[i]; }
	void clear()
	{
		for (int i = 0; i < w * h; i++) data[i] = T(0);
	}
	void copyTo(float* dst) const
	{
		for (int i = 0; i < w * h; i++) dst[i] = (float)data[i];
	}
	/* Takes src's size (reallocating when it differs) and pixels. */
	template <typename Ta>
	void set(const TImageData<Ta>& src)
	{
		if (!data || numPixels() != src.numPixels()) {
			free();
			w = src.w;
			h = src.h;
			if (src.data) data = new T[(size_t)w * h];
		}
		for (int i = 0; i < w * h && src.data; i++) data[i] = (T)src.data[i];
	}
	static TImageData alloc(int width, int height) { return TImageData(new T[(size_t)width * height], width, height); }
	void free()
	{
		delete[] data;
		data = nullptr;
	}
};

/* A float image that owns its pixels (new[] / delete[]). */
class CImageData : public TImageData<float> {
public:
	CImageData() {}
	CImageData(int width, int height) : TImageData<float>(new float[(size_t)width * height], width, height) {}
	CImageData(const CImageData& o) : TImageData<float>() { set(o); }
	CImageData(const TImageData<float>& o) : TImageData<float>() { set(o); }
	~CImageData() { free(); }
	CImageData& operator=(const CImageData& o)
	{
		if (this != &o) set(o);
		return *this;
	}
	CImageData& operator=(const TImageData<float>& o)
	{
		set(o);
		return *this;
	}
};

#endif
