// vx3_launch.h -- host entry points of the kernel translation units.
#ifndef VX3_LAUNCH_H
#define VX3_LAUNCH_H
#include <hip/hip_runtime.h>
#include "vx3_batch.h"

namespace vx3 {
struct StepOpts {
    bool poisson;      // any resident sim has a non-zero Poisson's ratio
    bool contacts;     // run the contact-force pass
    int nPairs;        // collision pairs in cpair
    int block;         // workgroup size of the element passes
};
void launchStep(const DevBatch<float>& b, hipStream_t s, const StepOpts& o);
void launchStep(const DevBatch<double>& b, hipStream_t s, const StepOpts& o);
void launchDt(const DevBatch<float>& b, hipStream_t s, int s0, int s1);
void launchDt(const DevBatch<double>& b, hipStream_t s, int s0, int s1);
}  // namespace vx3
#endif
