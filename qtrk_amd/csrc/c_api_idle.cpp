// c_api_idle.cpp -- the exported QTrkIsIdle.  The reference declares it `bool`
// (cputrack/qtrk_c_api.h:191), and QTrkDotNet imports it with C#'s default marshalling of `bool`, a
// 4-byte BOOL read from the whole return register (QTrkDotNet/QTrkDLL.cs).  A C++ `bool` return
// defines only the register's low byte; defined here with an `int` result of 0 or 1, it fills the
// register for such callers and reads the same for callers of the `bool` declaration.  This file
// does not include that declaration, which would clash with the definition.
extern "C" __attribute__((visibility("hidden"))) int qtrk_is_idle(void* qtrk);  // c_api.cpp

extern "C" __attribute__((visibility("default"))) int QTrkIsIdle(void* qtrk)
{
	return qtrk_is_idle(qtrk);
}
