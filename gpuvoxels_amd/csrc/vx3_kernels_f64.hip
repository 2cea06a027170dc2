// vx3_kernels_f64.hip -- stream-path kernels with double state.
#define VX_REAL double
#define VX_NS k_f64
#include "vx3_kernels.inl"
#include "vx3_resident.inl"
#include "vx3_launch.h"

#include "vx3_stream_launch.inl"

// resident kernel shapes: (block, voxels per thread, links per axis per thread); on
// equal voxel capacity the first listed is preferred (512x2x1: two waves per SIMD)
#define VX3_RES_SHAPES(X) X(512, 2, 1) X(256, 4, 2) X(256, 2, 1) X(256, 1, 1)
#define VX3_RES_TABLE_FN residentShapes_f64
#include "vx3_resident_launch.inl"

