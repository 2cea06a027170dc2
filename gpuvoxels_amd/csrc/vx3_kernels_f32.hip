// vx3_kernels_f32.hip -- stream-path kernels with float state.
#define VX_REAL float
#define VX_NS k_f32
#include "vx3_kernels.inl"
#include "vx3_resident.inl"
#include "vx3_launch.h"

namespace vx3 {

void launchStep(const DevBatch<float>& b, hipStream_t s, const StepOpts& o, hipEvent_t* ev) {
    const int bs = o.block > 0 ? o.block : 256;
    if (b.vHi <= 0) return;
    const int gv = (b.vHi + bs - 1) / bs, gl = (b.lHi + bs - 1) / bs;
#define VX3_EV(k, launch)                                   \
    do {                                                    \
        if (ev) hipEventRecord(ev[2 * (k)], s);             \
        launch;                                             \
        if (ev) hipEventRecord(ev[2 * (k) + 1], s);         \
    } while (0)
    if (o.poisson) VX3_EV(PASS_POISSON, hipLaunchKernelGGL(k_f32::k_poisson, dim3(gv), dim3(bs), 0, s, b));
    if (b.lHi > 0) VX3_EV(PASS_LINKS, hipLaunchKernelGGL(k_f32::k_links, dim3(gl), dim3(bs), 0, s, b));
    if (o.contacts && o.nPairs > 0)
        VX3_EV(PASS_CONTACTS, hipLaunchKernelGGL(k_f32::k_contacts, dim3((o.nPairs + 255) / 256), dim3(256), 0, s, b, o.nPairs));
    VX3_EV(PASS_VOXELS, hipLaunchKernelGGL(k_f32::k_voxels, dim3(gv), dim3(bs), 0, s, b));
    VX3_EV(PASS_CONTROL, hipLaunchKernelGGL(k_f32::k_control, dim3(b.nSims), dim3(256), 0, s, b));
#undef VX3_EV
}

void launchDt(const DevBatch<float>& b, hipStream_t s, int s0, int s1) {
    if (s1 > s0) hipLaunchKernelGGL(k_f32::k_dt, dim3(s1 - s0), dim3(256), 0, s, b, s0, s1);
}

}  // namespace vx3

// resident kernel shapes: (block, voxels per thread, links per axis per thread)
#define VX3_RES_SHAPES(X) X(1024, 1, 1) X(512, 2, 1) X(256, 4, 2) X(256, 2, 1) X(256, 1, 1)
#define VX3_RES_TABLE_FN residentShapes_f32
#include "vx3_resident_launch.inl"

