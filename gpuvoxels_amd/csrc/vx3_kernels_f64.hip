// vx3_kernels_f64.hip -- stream-path kernels with double state.
#define VX_REAL double
#define VX_NS k_f64
#include "vx3_kernels.inl"
#include "vx3_launch.h"

namespace vx3 {

void launchStep(const DevBatch<double>& b, hipStream_t s, const StepOpts& o) {
    const int bs = o.block > 0 ? o.block : 256;
    if (b.vHi <= 0) return;
    const int gv = (b.vHi + bs - 1) / bs, gl = (b.lHi + bs - 1) / bs;
    if (o.poisson) hipLaunchKernelGGL(k_f64::k_poisson, dim3(gv), dim3(bs), 0, s, b);
    if (b.lHi > 0) hipLaunchKernelGGL(k_f64::k_links, dim3(gl), dim3(bs), 0, s, b);
    if (o.contacts && o.nPairs > 0) hipLaunchKernelGGL(k_f64::k_contacts, dim3((o.nPairs + 255) / 256), dim3(256), 0, s, b, o.nPairs);
    hipLaunchKernelGGL(k_f64::k_voxels, dim3(gv), dim3(bs), 0, s, b);
    hipLaunchKernelGGL(k_f64::k_control, dim3(b.nSims), dim3(256), 0, s, b);
}

void launchDt(const DevBatch<double>& b, hipStream_t s, int s0, int s1) {
    if (s1 > s0) hipLaunchKernelGGL(k_f64::k_dt, dim3(s1 - s0), dim3(256), 0, s, b, s0, s1);
}

}  // namespace vx3
