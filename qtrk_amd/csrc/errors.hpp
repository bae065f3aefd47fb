// errors.hpp -- error plumbing for the C ABI: exceptions raised inside the
// library are caught at the boundary, printed to stderr and kept as the
// thread's last error (QTrkGetLastError).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <stdexcept>
#include <string>

namespace qtrk {

void set_last_error(const std::string& msg);

[[noreturn]] inline void fail(const std::string& msg) { throw std::runtime_error(msg); }

inline void hip_check(hipError_t e, const char* what)
{
	if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

// Runs f, converting any exception into the last error; returns false on failure.
template <typename F>
bool guard(const char* fn, F&& f)
{
	try {
		f();
		return true;
	} catch (const std::exception& e) {
		std::string m = std::string(fn) + ": " + e.what();
		fprintf(stderr, "qtrk: %s\n", m.c_str());
		set_last_error(m);
	} catch (...) {
		std::string m = std::string(fn) + ": unknown error";
		fprintf(stderr, "qtrk: %s\n", m.c_str());
		set_last_error(m);
	}
	return false;
}

}  // namespace qtrk

#define HIP_CHECK(x) ::qtrk::hip_check((x), #x)
// Body of a C-ABI function: exceptions become the last error; true on success.
#define GUARD(...) ::qtrk::guard(__func__, [&] { __VA_ARGS__; })
