// swd_internal.h -- shared declarations of libswd's translation units.
#pragma once
#include <cstdarg>
#include <cstddef>

// Records a formatted message as the calling thread's last error and returns code.
int swd_fail(int code, const char *fmt, ...);
