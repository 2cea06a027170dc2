// vx3_kernels_f32.hip -- stream-path kernels with float state.
#define VX_REAL float
#define VX_NS k_f32
#include "vx3_kernels.inl"
#include "vx3_resident.inl"
#include "vx3_launch.h"

#include "vx3_stream_launch.inl"

// resident kernel shapes: (block, voxels per thread, links per axis per thread)
#define VX3_RES_SHAPES(X) X(1024, 1, 1) X(512, 2, 1) X(256, 4, 2) X(256, 2, 1) X(256, 1, 1)
#define VX3_RES_TABLE_FN residentShapes_f32
#include "vx3_resident_launch.inl"

