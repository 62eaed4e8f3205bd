// odx_host.h -- host-side helpers shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "odx_device.h"

void odx_message(int err, const char *fmt, ...);
bool odx_check(hipError_t e, const char *what);
void odx_invert_pd3(const float *I /*3x4*/, float *out /*3x4*/);
// solver parameter vector (odxBatchCreate's params) -> step parameters and surface
void odx_params_from(const double *p, odx::StepParams &P, odx::Surface &S);
// body [nb][32] and geom [ng][16] records -> device SoA of D; allocations are appended to allocs
bool odx_state_upload(odx::DevWorlds &D, int nb, const float *bodies, int ng, const float *geoms,
                      std::vector<void *> &allocs);
bool odx_bodies_download(const odx::DevWorlds &D, int nb, float *out);
bool odx_bodies_upload(const odx::DevWorlds &D, int nb, const float *bodies);
// a material-pair surface table (float[nmat][nmat][16]) checked against the
// geoms' materials and converted for the device; false (with a message) if invalid
bool odx_surface_table(const char *who, int nmat, const float *table, const std::vector<int> &geom_mat,
                       std::vector<odx::Surface> &t);
