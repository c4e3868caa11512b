// Cell sort internals shared with the gravity step.
#pragma once

#include "pm_common.h"

namespace pm {
namespace sort {

int check_sort_args(int64_t n, const int32_t* mesh_shape, const int32_t* brick);
size_t workspace_bytes(int64_t n, const int32_t* mesh_shape, const int32_t* brick);
int sort_cells(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
               const int32_t* mesh_shape, float cell_size, const int32_t* brick, int32_t* perm,
               int32_t* bin_offsets, Arena& ws, hipStream_t s);

}  // namespace sort
}  // namespace pm
