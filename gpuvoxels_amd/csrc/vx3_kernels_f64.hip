// vx3_kernels_f64.hip -- stream-path kernels with double state.
#define VX_REAL double
#define VX_NS k_f64
#include "vx3_kernels.inl"
#include "vx3_resident.inl"
#include "vx3_launch.h"

namespace vx3 {

void launchStep(const DevBatch<double>& b, hipStream_t s, const StepOpts& o, hipEvent_t* ev) {
    const int bs = o.block > 0 ? o.block : 256;
    if (b.vHi <= 0) return;
    const int gv = (b.vHi + bs - 1) / bs, gl = (b.lHi + bs - 1) / bs;
#define VX3_EV(k, launch)                                   \
    do {                                                    \
        if (ev) hipEventRecord(ev[2 * (k)], s);             \
        launch;                                             \
        if (ev) hipEventRecord(ev[2 * (k) + 1], s);         \
    } while (0)
    if (o.poisson) VX3_EV(PASS_POISSON, hipLaunchKernelGGL(k_f64::k_poisson, dim3(gv), dim3(bs), 0, s, b));
    if (b.lHi > 0) VX3_EV(PASS_LINKS, hipLaunchKernelGGL(k_f64::k_links, dim3(gl), dim3(bs), 0, s, b));
    if (o.contacts && o.nPairs > 0)
        VX3_EV(PASS_CONTACTS, hipLaunchKernelGGL(k_f64::k_contacts, dim3((o.nPairs + 255) / 256), dim3(256), 0, s, b, o.nPairs));
    VX3_EV(PASS_VOXELS, hipLaunchKernelGGL(k_f64::k_voxels, dim3(gv), dim3(bs), 0, s, b));
    VX3_EV(PASS_CONTROL, hipLaunchKernelGGL(k_f64::k_control, dim3(b.nSims), dim3(256), 0, s, b));
#undef VX3_EV
}

int residentShapes_f64(ResidentShapeInfo* out, int cap) {
    static const ResidentShapeInfo shapes[] = {
        {256, 2, 4, k_f64::ResidentShape<256, 2, 4>::bytes},
        {256, 1, 3, k_f64::ResidentShape<256, 1, 3>::bytes},
        {256, 1, 2, k_f64::ResidentShape<256, 1, 2>::bytes},
    };
    const int n = (int)(sizeof(shapes) / sizeof(shapes[0]));
    for (int i = 0; i < n && i < cap; i++) out[i] = shapes[i];
    return n;
}

hipError_t launchResident(const DevBatch<double>& b, hipStream_t s, int shape, const int* simList, int nSims, int nSteps, hipEvent_t* ev) {
    if (nSims <= 0 || nSteps <= 0) return hipSuccess;
    switch (shape) {
    case 0: {
        auto kern = k_f64::k_resident<256, 2, 4>;
        const size_t lds = k_f64::ResidentShape<256, 2, 4>::bytes;
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        if (ev) hipEventRecord(ev[0], s);
        hipLaunchKernelGGL(kern, dim3(nSims), dim3(256), lds, s, b, simList, nSteps);
        if (ev) hipEventRecord(ev[1], s);
        return hipGetLastError();
    }
    case 1: {
        auto kern = k_f64::k_resident<256, 1, 3>;
        const size_t lds = k_f64::ResidentShape<256, 1, 3>::bytes;
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        if (ev) hipEventRecord(ev[0], s);
        hipLaunchKernelGGL(kern, dim3(nSims), dim3(256), lds, s, b, simList, nSteps);
        if (ev) hipEventRecord(ev[1], s);
        return hipGetLastError();
    }
    case 2: {
        auto kern = k_f64::k_resident<256, 1, 2>;
        const size_t lds = k_f64::ResidentShape<256, 1, 2>::bytes;
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        if (ev) hipEventRecord(ev[0], s);
        hipLaunchKernelGGL(kern, dim3(nSims), dim3(256), lds, s, b, simList, nSteps);
        if (ev) hipEventRecord(ev[1], s);
        return hipGetLastError();
    }
    default: return hipErrorInvalidValue;
    }
}

void launchDt(const DevBatch<double>& b, hipStream_t s, int s0, int s1) {
    if (s1 > s0) hipLaunchKernelGGL(k_f64::k_dt, dim3(s1 - s0), dim3(256), 0, s, b, s0, s1);
}

}  // namespace vx3
