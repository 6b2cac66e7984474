// nw_amg.h -- smoothed-aggregation AMG preconditioner (device side).
#pragma once

#include "nw_amg_setup.h"
#include "nw_internal.h"

namespace nw {

struct AmgDevice;

// Builds the hierarchy from the context's current 1-dof matrix (host setup)
// and uploads it in FP64.
int amg_setup(Ctx* c, AmgDevice** out);
void amg_free(AmgDevice* a);
int amg_levels(const AmgDevice* a);
// wall time of amg_setup: matrix download, host build, device upload
double amg_setup_seconds(const AmgDevice* a);
// rows and stored entries of level `level` (0 = the fine matrix)
int amg_level_info(const AmgDevice* a, int level, int64_t* rows, int64_t* nnz);
// Makes the hierarchy in type T on first use (FP32: narrowed from the FP64
// one on the device; *seconds, if given, receives the time that took, 0
// when it existed) and refreshes the fine-level values from the matrix.
template <typename T>
int amg_prepare(Ctx* c, AmgDevice* a, double* seconds = nullptr);
// x = M^-1 b (one V-cycle); x and b are fine-level vectors, x != b.
template <typename T>
int amg_apply(AmgDevice* a, const T* b, T* x, hipStream_t st);

}  // namespace nw
