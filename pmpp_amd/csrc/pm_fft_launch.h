// Host-side launchers of the FFT passes (defined in pm_fft.hip), shared with pm_gravity.hip.
#pragma once

#include "pm_fft.h"

namespace pm {
namespace fft {

bool supported_shape(const int32_t* shape);
int launch_cols(int n, int mode, const ColPass& cp, int batches, const SpecTables& tb, int dax,
                hipStream_t s);
int launch_z(int nz, bool c2r, const ZPass& zp, const SpecTables& tb, hipStream_t s);

// Spectral tables: 3 * (nx + ny + nz/2+1) floats, laid out per axis as k | mask | negk.
size_t spec_table_floats(const int32_t* shape);
SpecTables spec_scalars(double cell_size, double ptcl_spacing, int mask_active);
SpecTables spec_tables_view(const float* base, const int32_t* shape, const SpecTables& scalars);
int launch_spec_tables(float* out, const int32_t* shape, const SpecTables& scalars, hipStream_t s);
int launch_spectral_multiply(const float2* in, float2* out, const int32_t* shape, const SpecTables& tb,
                             int axis, int adjoint, hipStream_t s);

int rfft3d_pitched(const float* in, float2* spec, const int32_t* shape, int64_t pitch, float shift,
                   float scale, const SpecTables& tb, hipStream_t s);
int irfft3d_pitched(float2* spec, float* out, const int32_t* shape, int64_t pitch, const SpecTables& tb,
                    hipStream_t s);

}  // namespace fft
}  // namespace pm
