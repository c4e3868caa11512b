// Cell sort internals shared with the gravity step.
#pragma once

#include "pm_common.h"

namespace pm {
namespace sort {

int check_sort_args(int64_t n, const int32_t* mesh_shape, const int32_t* brick);
size_t workspace_bytes(int64_t n, const int32_t* mesh_shape, const int32_t* brick);
int sort_cells(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n,
               const int32_t* mesh_shape, float cell_size, const int32_t* brick, int32_t* perm,
               int32_t* bin_offsets, int16_t* pmid_s, float* disp_s, Arena& ws, hipStream_t s);

// Unstable bucketing by the same key (gravity step): bin_offsets (num_bricks + 2 entries, the
// last bucket holding invalid particles), perm and the payload gathered into bucket order.
// ptcl_grid (may be null) is the Lagrangian grid the particles were generated on, a hint
// that selects grid-block chunks (speed only).
size_t bucket_workspace_bytes(const int32_t* mesh_shape, const int32_t* brick);
int bucket_cells(const int16_t* pmid, const float* disp, const uint8_t* valid, int64_t n, const int32_t* mesh_shape,
                 float cell_size, const int32_t* brick, const int32_t* ptcl_grid, int32_t* perm, int32_t* bin_offsets,
                 int16_t* pmid_s, float* disp_s, uint8_t* valid_s, Arena& ws, hipStream_t s);

// pm_init: resolve the current device's address of the allocation-free binning scratch.
int init_noalloc_scratch();

}  // namespace sort
}  // namespace pm
