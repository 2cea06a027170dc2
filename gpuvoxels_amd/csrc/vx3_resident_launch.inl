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
        hipLaunchKernelGGL(kern, dim3(nSims), dim3(B), lds, s, b, simList, nSteps, TileArgs{});                              \
        if (ev) hipEventRecord(ev[1], s);                                                                        \
        return hipGetLastError();                                                                                \
    }
    VX3_RES_SHAPES(VX3_CASE)
#undef VX3_CASE
    return hipErrorInvalidValue;
}

// Tiled launches: one workgroup per tile of the simulations in ta.tiles[0, nTiles), all
// resident at once (a cooperative launch), after the simulations' sync words
// ta.sync[0, nSync) (divergence words and arrival flags) are cleared; the repair launch
// keeps them (it reads the divergence words, and its flags count from their own base).
hipError_t launchTiled(const DevBatch<VX_REAL>& b, hipStream_t s, int shape, const TileArgs& ta, int nTiles, int nSync,
                       int nSteps, hipEvent_t* ev) {
    if (nTiles <= 0 || nSteps <= 0) return hipSuccess;
    int i = 0;
#define VX3_CASE(B, V, L)                                                                                        \
    if (shape == i++) {                                                                                          \
        auto kern = VX_NS::k_resident<B, V, L, false, true>;                                                     \
        const size_t lds = VX_NS::ResidentShape<B, V, L, false, true>::bytes;                                    \
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        if (e != hipSuccess) return e;                                                                           \
        DevBatch<VX_REAL> bb = b;                                                                                \
        const int* none = nullptr;                                                                               \
        int steps = nSteps;                                                                                      \
        TileArgs t = ta;                                                                                         \
        void* args[] = {(void*)&bb, (void*)&none, (void*)&steps, (void*)&t};                                     \
        for (int pass = 0; pass < 2; pass++) {   /* the launch, then its repair launch */                        \
            t.repair = pass;                                                                                     \
            if (!pass)                                                                                           \
                hipLaunchKernelGGL(VX_NS::k_fill_ints, dim3((nSync + 255) / 256), dim3(256), 0, s, ta.sync, nSync, 1, 0); \
            if (ev && !pass) hipEventRecord(ev[0], s);                                                           \
            e = hipLaunchCooperativeKernel((const void*)kern, dim3(nTiles), dim3(B), args, (unsigned)lds, s);    \
            if (ev && !pass) hipEventRecord(ev[1], s);                                                           \
            if (e != hipSuccess) return e;                                                                       \
        }                                                                                                        \
        return hipGetLastError();                                                                                \
    }
    VX3_RES_SHAPES(VX3_CASE)
#undef VX3_CASE
    return hipErrorInvalidValue;
}

// workgroups of shape `shape`'s tiled kernel that one compute unit holds at once
hipError_t tiledOccupancy(const DevBatch<VX_REAL>&, int shape, int* perCU) {
    int i = 0;
#define VX3_CASE(B, V, L)                                                                                        \
    if (shape == i++) {                                                                                          \
        auto kern = VX_NS::k_resident<B, V, L, false, true>;                                                     \
        const size_t lds = VX_NS::ResidentShape<B, V, L, false, true>::bytes;                                    \
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        if (e != hipSuccess) return e;                                                                           \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(perCU, (const void*)kern, B, lds);                   \
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
