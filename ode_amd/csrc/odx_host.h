// odx_host.h -- host-side helpers shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

void odx_message(int err, const char *fmt, ...);
bool odx_check(hipError_t e, const char *what);
void odx_invert_pd3(const float *I /*3x4*/, float *out /*3x4*/);
