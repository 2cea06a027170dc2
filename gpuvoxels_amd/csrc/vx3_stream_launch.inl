// vx3_stream_launch.inl -- host launchers of one step of the stream path and of the
// recommended-dt pass, for the precision of the including unit (VX_REAL, VX_NS).

namespace vx3 {

namespace {
void launchCollisions(const DevBatch<VX_REAL>& b, hipStream_t s, int bs, bool topo) {
    const int gv = (b.vHi + bs - 1) / bs;
    if (topo) {   // nearby lists of simulations whose topology changed in the last step
        const int gn = (b.vHi + VX3_NEAR_BS - 1) / VX3_NEAR_BS;
        hipLaunchKernelGGL(VX_NS::k_near_count, dim3(gn), dim3(VX3_NEAR_BS), 0, s, b);
        hipLaunchKernelGGL(VX_NS::k_near_emit, dim3(gn), dim3(VX3_NEAR_BS), 0, s, b);
    }
    // the staleness pass (colStalePass) and the clearing of the bucket counts ran with the
    // link pass (k_links_stale)
    hipLaunchKernelGGL(VX_NS::k_col_insert, dim3(gv), dim3(bs), 0, s, b);
    if (b.hashTot > 0) hipLaunchKernelGGL(VX_NS::k_col_bscan, dim3(b.hashTot / VX3_COL_HBLOCK), dim3(256), 0, s, b);
    hipLaunchKernelGGL(VX_NS::k_col_place, dim3(gv), dim3(bs), 0, s, b);
    if (b.cellHi > 0) hipLaunchKernelGGL(VX_NS::k_col_list, dim3((b.cellHi + 255) / 256), dim3(256), 0, s, b);
}
}  // namespace

void launchStep(const DevBatch<VX_REAL>& b, hipStream_t s, const StepOpts& o, hipEvent_t* ev) {
    const int bs = o.block > 0 ? o.block : 256;
    if (b.vHi <= 0) return;
    const int gv = (b.vHi + bs - 1) / bs, gl = (b.lHi + bs - 1) / bs;
#define VX3_EV(k, launch)                                   \
    do {                                                    \
        if (ev) hipEventRecord(ev[2 * (k)], s);             \
        launch;                                             \
        if (ev) hipEventRecord(ev[2 * (k) + 1], s);         \
    } while (0)
    if (o.prestep) hipLaunchKernelGGL(VX_NS::k_prestep, dim3(gv), dim3(bs), 0, s, b, 0);
    if (o.attach) hipLaunchKernelGGL(VX_NS::k_attach, dim3(b.nSims), dim3(256), 0, s, b);
    // The collision lists are rebuilt between the link and voxel passes, as the reference
    // does (CVoxelyze::doTimeStep, Voxelyze.cpp:251-284): the link pass moves no voxel, and
    // nearby lists regenerated after a topology change follow the links this step's link
    // pass left (its detaches included).  A simulation whose table proves too small stands
    // still after its link pass (SC_LINKS_DONE) and resumes there once the table is grown.
    if (o.contacts)   // with the collisions' staleness pass, one thread per link and per voxel
        VX3_EV(PASS_LINKS, hipLaunchKernelGGL(VX_NS::k_links_stale, dim3(gl > gv ? gl : gv), dim3(bs), 0, s, b));
    else if (b.lHi > 0)
        VX3_EV(PASS_LINKS, hipLaunchKernelGGL(VX_NS::k_links, dim3(gl), dim3(bs), 0, s, b));
    if (o.contacts) VX3_EV(PASS_CONTACTS, launchCollisions(b, s, bs, o.topo));
    if (o.contacts && o.extras) hipLaunchKernelGGL(VX_NS::k_hits, dim3(gv), dim3(bs), 0, s, b);
    if (o.extras) hipLaunchKernelGGL(VX_NS::k_prestep, dim3(gv), dim3(bs), 0, s, b, 1);
    if (o.extras) hipLaunchKernelGGL(VX_NS::k_voxels_extras, dim3(gv), dim3(bs), 0, s, b);
    VX3_EV(PASS_VOXELS, hipLaunchKernelGGL(VX_NS::k_voxels, dim3(gv), dim3(bs), 0, s, b));
    if (o.signals) hipLaunchKernelGGL(VX_NS::k_signals, dim3(gv), dim3(bs), 0, s, b);
    if (o.attach) hipLaunchKernelGGL(VX_NS::k_prestep, dim3(gv), dim3(bs), 0, s, b, 2);
    VX3_EV(PASS_CONTROL, hipLaunchKernelGGL(VX_NS::k_control, dim3(b.nSims), dim3(256), 0, s, b));
#undef VX3_EV
}

void launchCopyLoads(const DevBatch<VX_REAL>& b, hipStream_t s, const void* src, size_t srcFs, void* dst, size_t dstFs, int nv) {
    if (nv > 0)
        hipLaunchKernelGGL(VX_NS::k_copy_loads, dim3((nv + 255) / 256), dim3(256), 0, s, (const VX_REAL*)src, srcFs,
                           (VX_REAL*)dst, dstFs, nv);
}

void launchDt(const DevBatch<VX_REAL>& b, hipStream_t s, int s0, int s1) {
    if (s1 > s0) hipLaunchKernelGGL(VX_NS::k_dt, dim3(s1 - s0), dim3(256), 0, s, b, s0, s1);
}

}  // namespace vx3
