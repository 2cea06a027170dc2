// vx3_resident_launch.inl -- host-side table and launcher of the resident kernel
// shapes of one precision.  The including unit defines VX3_RES_SHAPES(X), a list of
// X(block, voxels per thread, links per axis per thread), and VX3_RES_TABLE_FN.

namespace vx3 {

static_assert(VX3_RES_VMATS == RES_MAX_VMATS && VX3_RES_LMATS == RES_MAX_LMATS && VX3_RES_PHASES == RES_MAX_PHASES,
              "resident kernel table sizes and admission limits agree");

int VX3_RES_TABLE_FN(ResidentShapeInfo* out, int cap) {
#define VX3_ROW(B, V, L) {B, V, L, VX_NS::ResidentShape<B, V, L>::bytes, VX_NS::ResidentShape<B, V, L, true>::bytes},
    static const ResidentShapeInfo shapes[] = {VX3_RES_SHAPES(VX3_ROW)};
#undef VX3_ROW
    const int n = (int)(sizeof(shapes) / sizeof(shapes[0]));
    for (int i = 0; i < n && i < cap; i++) out[i] = shapes[i];
    return n;
}

template <bool COL>
hipError_t launchResidentAs(const DevBatch<VX_REAL>& b, hipStream_t s, int shape, const int* simList, int nSims, int nSteps,
                            hipEvent_t* ev) {
    int i = 0;
#define VX3_CASE(B, V, L)                                                                                        \
    if (shape == i++) {                                                                                          \
        auto kern = VX_NS::k_resident<B, V, L, COL>;                                                             \
        const size_t lds = VX_NS::ResidentShape<B, V, L, COL>::bytes;                                            \
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        if (e != hipSuccess) return e;                                                                           \
        if (ev) hipEventRecord(ev[0], s);                                                                        \
        hipLaunchKernelGGL(kern, dim3(nSims), dim3(B), lds, s, b, simList, nSteps);                              \
        if (ev) hipEventRecord(ev[1], s);                                                                        \
        return hipGetLastError();                                                                                \
    }
    VX3_RES_SHAPES(VX3_CASE)
#undef VX3_CASE
    return hipErrorInvalidValue;
}

hipError_t launchResident(const DevBatch<VX_REAL>& b, hipStream_t s, int shape, bool collisions, const int* simList, int nSims,
                          int nSteps, hipEvent_t* ev) {
    if (nSims <= 0 || nSteps <= 0) return hipSuccess;
    return collisions ? launchResidentAs<true>(b, s, shape, simList, nSims, nSteps, ev)
                      : launchResidentAs<false>(b, s, shape, simList, nSims, nSteps, ev);
}

}  // namespace vx3
