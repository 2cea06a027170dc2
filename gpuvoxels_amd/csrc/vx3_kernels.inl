// vx3_kernels.inl -- stream-path kernels, instantiated for one state precision
// by the including translation unit (VX_REAL = float or double, VX_SUFFIX = f32/f64).
//
// One step of every running simulation of the batch (the body of
// VX3_VoxelyzeKernel::doTimeStep, Voxelyze3/src/VX3/VX3_VoxelyzeKernel.cu:237-359,
// with the CPU engine's per-step order, Voxelyze3/src/old/Voxelyze.cpp:251-284):
//   k_links     beam element of every link, writes the two end loads into the
//               (direction, voxel) force slots; flags divergence
//   k_col_*     collision neighbour lists, rebuilt when stale (collisions only)
//   k_voxels    gathers the six slots, adds contacts, integrates, writes next step's
//               thermal factor and Poisson strain (ν ≠ 0 only)
//   k_control   per simulation: step count, divergence, periodic FP64 center of
//               mass, time advance, stop condition
#include "vx3_batch.h"
#include "vx3_physics.h"
#include "vx3_powf.h"

namespace vx3 {
namespace VX_NS {

typedef VX_REAL real;

// a stream-path simulation that takes this step (not finished, not on the resident
// path, not standing still for a larger collision table)
__device__ __forceinline__ bool simRunning(const DevBatch<real>& b, int s) {
    if (s < 0 || b.sims[s].state != 0) return false;
    const int f = b.sims[s].flags;
    if (b.solo >= 0) return s == b.solo && !(f & SC_COL_GROW);
    if (b.solo == -2) return !(f & SC_COL_GROW);   // every simulation, resident ones included
    return !(f & (SC_RESIDENT | SC_COL_GROW));
}

// Precision-dependent transcendentals.  The FP64 state path rounds like the host C
// library (pow of float arguments evaluated in double and rounded once; sin in double);
// the FP32 state path uses single-precision functions on a double-reduced argument.
__device__ __forceinline__ float poissonValue(float base, float nu) {   // pow(base, -nu) - 1
    if constexpr (sizeof(real) == 8) {
        float p;
        if (powNearOne(base, -nu, p)) return p - 1.0f;   // same float as the line below, cheaper
        return (float)pow((double)base, (double)(-nu)) - 1.0f;
    } else return powf(base, -nu) - 1.0f;
}
// VX3 actuation temperature at time tn (VX3_VoxelyzeKernel.cu:625-650): A*sin(2*3.1415926f*(t/Tp + phase)),
// contraction only unless expansion is enabled
__device__ __forceinline__ float actuationTemp(double amp, double period, double phase, double tn, bool expansion) {
    const double arg = (double)(2 * 3.1415926f) * (tn / period + phase);
    if constexpr (sizeof(real) == 8) {
        double T = amp * sin(arg);
        if (!expansion && T > 0) T = 0;
        return (float)T;
    } else {
        const double twoPi = 6.283185307179586;
        const double r = arg - floor(arg * (1.0 / twoPi)) * twoPi;
        float T = (float)amp * sinf((float)r);
        if (!expansion && T > 0) T = 0;
        return T;
    }
}

// CVX_Voxel::transverseArea / transverseStrainSum (VX_Voxel.cpp:345-373) from the cached
// Poisson strain ps = (x, y, z) of the voxel
__device__ __forceinline__ void transverseOf(const float* ps, int axis, float size, bool nuZero, float& area, float& sum) {
    if (nuZero) { area = size * size; sum = 0.0f; return; }
    float sx = ps[0], sy = ps[1], sz = ps[2];
    real s2 = (real)(size * size);
    if (axis == 0) { area = (float)(s2 * ((real)1 + (real)sy) * ((real)1 + (real)sz)); sum = sy + sz; }
    else if (axis == 1) { area = (float)(s2 * ((real)1 + (real)sx) * ((real)1 + (real)sz)); sum = sx + sz; }
    else { area = (float)(s2 * ((real)1 + (real)sx) * ((real)1 + (real)sy)); sum = sx + sy; }
}

// Poisson strain of voxel v for the next step, CVX_Voxel::strain(true)
// (VX_Voxel.cpp:299-333), from the end strains this step's link pass left in the
// slots (the value the next step's link pass would compute lazily from the same
// strains).  Called at the end of the voxel pass.
__device__ __forceinline__ void poissonStrain(const DevBatch<real>& b, int v) {
    const int fl = b.vflags[v];
    float st[3] = {0, 0, 0};
    int n[3] = {0, 0, 0};
    for (int d = 0; d < 6; d++) {
        if (fl & (1 << (VF_LINKMASK_SHIFT + d))) {
            st[d >> 1] += b.es[d * b.Vcap + v];
            n[d >> 1]++;
        }
    }
    const int e = b.vext[v];
    const ExtDev* ext = e >= 0 ? &b.extTab[e] : nullptr;
    bool tension[3];
    for (int i = 0; i < 3; i++) {
        if (n[i] == 2) st[i] *= 0.5f;
        tension[i] = (n[i] == 2) || (ext && (n[i] == 1 && ((ext->dofFixed & (1 << i)) || ext->force[i] != 0)));
    }
    if (!(tension[0] && tension[1] && tension[2])) {
        float add = 0;
        for (int i = 0; i < 3; i++) if (tension[i]) add += st[i];
        // pow of float arguments: evaluated in double and rounded once, which gives the
        // correctly rounded single-precision result the host C library returns
        const float value = poissonValue(1.0f + add, b.vmatTab[b.vmat[v]].nu);
        for (int i = 0; i < 3; i++) if (!tension[i]) st[i] = value;
    }
    b.psx[v] = st[0]; b.psy[v] = st[1]; b.psz[v] = st[2];
}

__device__ __forceinline__ void linkPass(const DevBatch<real>& b, int l) {
    if (l >= b.lHi) return;
    // The loads come in three rounds of independent requests, so that a link waits on
    // memory three times rather than once per table it visits: the link's own indices,
    // then its simulation's control block, then everything the force needs (link state,
    // both end voxels' pose, thermal factor and Poisson strain, their materials).
    const int s = b.lsim[l];
    const int vn = b.lneg[l], vp = b.lpos[l];
    const int meta = b.lmeta[l];
    const int lf0 = b.lflags[l];
    if (!simRunning(b, s)) return;
    const SimCtl& sc = b.sims[s];
    if (sc.flags & SC_LINKS_DONE) return;   // resuming after this step's link pass
    if (lf0 & LF_GONE) return;              // detached (VX3_VoxelyzeKernel.cu:574)
    if (lf0 & LF_FRESH) {                   // made by attach after this step's link pass in the reference
        b.lflags[l] = lf0 & ~LF_FRESH;
        return;
    }
    const bool poisson = (sc.flags & SC_POISSON) != 0;
    const int vmn = b.vmat[vn], vmp = b.vmat[vp];
    LinkState<real> st;
    st.pos2 = V3<real>(b.q2x[l], b.q2y[l], b.q2z[l]);
    st.a1v = V3<real>(b.a1x[l], b.a1y[l], b.a1z[l]);
    st.a2v = V3<real>(b.a2x[l], b.a2y[l], b.a2z[l]);
    st.strain = b.strain[l];
    st.maxStrain = b.maxStrain[l];
    st.strainOffset = b.strainOffset[l];
    st.flags = lf0;
    const V3<real> pN(b.px[vn], b.py[vn], b.pz[vn]), pP(b.px[vp], b.py[vp], b.pz[vp]);
    const Q4<real> oN(b.ow[vn], b.ox[vn], b.oy[vn], b.oz[vn]), oP(b.ow[vp], b.ox[vp], b.oy[vp], b.oz[vp]);
    const float tfN = b.tfac[vn], tfP = b.tfac[vp];
    float psN[3] = {0, 0, 0}, psP[3] = {0, 0, 0};
    if (poisson) {
        psN[0] = b.psx[vn]; psN[1] = b.psy[vn]; psN[2] = b.psz[vn];
        psP[0] = b.psx[vp]; psP[1] = b.psy[vp]; psP[2] = b.psz[vp];
    }
    const int axis = meta & 3;
    const LinkMatDev& mat = b.lmatTab[meta >> 2];
    const VoxMatDev& mn = b.vmatTab[vmn];
    const VoxMatDev& mp = b.vmatTab[vmp];
    if (mn.fixed && mp.fixed) return;   // VX3_VoxelyzeKernel.cu:572
    const real nom = (real)sc.voxSize;
    // CVX_Link::updateRestLength: 0.5*(baseSize_neg + baseSize_pos)
    const real restLength = (real)0.5 * (nom * (real)tfN + nom * (real)tfP);
    float area, tss;
    const float size = (float)sc.voxSize;
    if (mat.nu != 0.0f) {
        float an, sn, ap, sp;
        transverseOf(psN, axis, size, mn.nu == 0.0f, an, sn);
        transverseOf(psP, axis, size, mp.nu == 0.0f, ap, sp);
        area = 0.5f * (an + ap);
        tss = 0.5f * (sn + sp);
    } else {
        area = 0.5f * (size * size + size * size);
        tss = 0.0f;
    }
    // damping multipliers 2*sqrt(m)*zeta/previousDt (VX_Voxel.h:131), previousDt being the dt of the
    // voxel's last update; material-fixed voxels never step
    const float dmN = 2 * mn.sqrtMass * mn.zetaInternal / (mn.fixed ? 0.0f : sc.prevDt);
    const float dmP = 2 * mp.sqrtMass * mp.zetaInternal / (mp.fixed ? 0.0f : sc.prevDt);
    V3<real> fN, mN, fP, mP;
    linkForces<real>(mat, axis, pN, oN, pP, oP, restLength, area, tss, mat.nu != 0.0f, dmN, dmP, st, fN, mN, fP, mP);
    if ((sc.flags & SC_ATTACH) && b.lnew[l] > 0) {   // isNewLink (VX3_Link.cu:206-213)
        const real g = (real)0.01;
        fN = fN.mul(g); mN = mN.mul(g); fP = fP.mul(g); mP = mP.mul(g);
        b.lnew[l] -= 1;
    }
    b.q2x[l] = st.pos2.x; b.q2y[l] = st.pos2.y; b.q2z[l] = st.pos2.z;
    b.a1x[l] = st.a1v.x; b.a1y[l] = st.a1v.y; b.a1z[l] = st.a1v.z;
    b.a2x[l] = st.a2v.x; b.a2y[l] = st.a2v.y; b.a2z[l] = st.a2v.z;
    b.strain[l] = st.strain;
    b.maxStrain[l] = st.maxStrain;
    b.strainOffset[l] = st.strainOffset;
    b.lflags[l] = st.flags;
    // slot of the negative end: the link lies in direction 2*axis (X_POS/Y_POS/Z_POS) from it
    const int dn = 2 * axis, dp = 2 * axis + 1;
    real* F = b.F;
    const size_t V = (size_t)b.Vcap, FS = b.Fs;
    F[(dn * 6 + 0) * FS + vn] = fN.x; F[(dn * 6 + 1) * FS + vn] = fN.y; F[(dn * 6 + 2) * FS + vn] = fN.z;
    F[(dn * 6 + 3) * FS + vn] = mN.x; F[(dn * 6 + 4) * FS + vn] = mN.y; F[(dn * 6 + 5) * FS + vn] = mN.z;
    F[(dp * 6 + 0) * FS + vp] = fP.x; F[(dp * 6 + 1) * FS + vp] = fP.y; F[(dp * 6 + 2) * FS + vp] = fP.z;
    F[(dp * 6 + 3) * FS + vp] = mP.x; F[(dp * 6 + 4) * FS + vp] = mP.y; F[(dp * 6 + 5) * FS + vp] = mP.z;
    if (sc.flags & SC_POISSON) {
        // CVX_Link::axialStrain(positiveEnd) (VX_Link.cpp:115-118)
        const float r = mat.strainRatio;
        b.es[dn * V + vn] = 2.0f * st.strain / (1.0f + r);
        b.es[dp * V + vp] = 2.0f * st.strain * r / (1.0f + r);
    }
    if (st.strain > 100) atomicOr(&b.sims[s].flags, SC_DIVERGED_PENDING);   // Voxelyze.cpp:265
    // gpu_update_detach (VX3_VoxelyzeKernel.cu:946-967): a failed link leaves both voxels
    // before their update; their surface state follows at the next step's start
    if ((sc.flags & SC_DETACH) && mat.epsilonFail != -1.0f && st.maxStrain > mat.epsilonFail) {
        b.lflags[l] = st.flags | LF_GONE;
        if (sc.flags & SC_TOPOCOL) atomicOr(&b.sims[s].flags, SC_TOPO_CHG);
        atomicAnd(&b.vflags[vn], ~(1 << (VF_LINKMASK_SHIFT + dn)));
        atomicAnd(&b.vflags[vp], ~(1 << (VF_LINKMASK_SHIFT + dp)));
    }
}

__global__ void __launch_bounds__(256) k_links(DevBatch<real> b) { linkPass(b, blockIdx.x * blockDim.x + threadIdx.x); }

// ---- collisions (CVoxelyze::updateCollisions / regenerateCollisions, Voxelyze.cpp:671-751)
//
// Instead of the CPU engine's all-pairs regeneration, the surface voxels of every
// simulation being rebuilt are sorted into cells (edge = watch radius) by a counting
// sort over a per-simulation hash of the cell coordinates, and each voxel finds its
// partners in the 27 neighbouring cells.  The partner criterion is the CPU one: both
// voxels on the surface, centre distance^2 <= watchRadius^2, and the partner not
// "nearby" (within the link-hop radius, CVX_Voxel::generateNearby).  Each voxel keeps
// its partners in ascending index order -- the order in which CVX_Voxel::force walks
// the colWatch list built by the pair loop -- and computes the contact force of each
// pair itself, in the voxel pass.
//
// Sort: k_col_insert counts the voxels of each hash bucket (and each voxel's rank in
// its bucket), k_col_bscan turns the counts into offsets within blocks of
// VX3_COL_HBLOCK buckets and, in the simulation's last block to finish, into
// per-simulation offsets of those blocks, and
// k_col_place scatters the voxels into the simulation's cell-ordered range of cellVox.
// No bucket has a capacity.  The partner lists are counted, scanned and emitted; if a
// simulation's partners exceed its neighbour table, it sets SC_COL_GROW and stands
// still after its link pass (SC_LINKS_DONE: no later pass changes its state) until the
// host enlarges the table; its next step then resumes after the link pass.

__device__ __forceinline__ bool colVoxel(const DevBatch<real>& b, int v, int& s) {
    s = b.vsim[v];
    if (!simRunning(b, s)) return false;
    const int f = b.sims[s].flags;
    if (!(f & SC_COLLISIONS) || (f & (SC_DIVERGED_PENDING | SC_COL_GROW))) return false;
    return !(b.vflags[v] & VF_INTERIOR);
}

__device__ __forceinline__ void colCell(const DevBatch<real>& b, int v, float cell, int c[3]) {
    c[0] = (int)floor((double)b.px[v] / (double)cell);
    c[1] = (int)floor((double)b.py[v] / (double)cell);
    c[2] = (int)floor((double)b.pz[v] / (double)cell);
}
__device__ __forceinline__ unsigned colHash(int x, int y, int z) {
    return ((unsigned)x * 73856093u) ^ ((unsigned)y * 19349663u) ^ ((unsigned)z * 83492791u);
}

__constant__ NearBitTable kNearBits = NearBitTable();

// whether global voxel u is in global voxel v's nearby set: its bit in v's mask (SC_NEARMASK),
// else a binary search of v's sorted nearby list (sim-local indices)
__device__ __forceinline__ bool nearbyOf(const DevBatch<real>& b, const SimCtl& sc, int v, int u) {
    if (sc.flags & SC_NEARMASK) {
        const int dx = b.vijk[3 * u] - b.vijk[3 * v], dy = b.vijk[3 * u + 1] - b.vijk[3 * v + 1],
                  dz = b.vijk[3 * u + 2] - b.vijk[3 * v + 2];
        const int bit = kNearBits.bit[(dx + 5) * 121 + (dy + 5) * 11 + (dz + 5)];   // |d| <= 5: checked by the caller
        return bit >= 0 && ((b.nearMask[(size_t)(bit >> 5) * b.Vcap + v] >> (bit & 31)) & 1u);
    }
    const int local = u - sc.v0;
    int lo = b.nearStart[v], hi = lo + b.nearCount[v];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int x = b.nearList[mid];
        if (x == local) return true;
        if (x < local) lo = mid + 1; else hi = mid;
    }
    return false;
}

// the CPU pair criterion between two surface voxels of one simulation
__device__ __forceinline__ bool colWatched(const DevBatch<real>& b, const SimCtl& sc, int v, int u) {
    const real dx = b.px[v] - b.px[u], dy = b.py[v] - b.py[u], dz = b.pz[v] - b.pz[u];
    if (dx * dx + dy * dy + dz * dz > (real)sc.colThreshSq) return false;
    // nearby: within the link-hop radius; graph distance >= lattice L1 distance, so only
    // lattice-close voxels can be nearby
    const int li = abs(b.vijk[3 * v] - b.vijk[3 * u]) + abs(b.vijk[3 * v + 1] - b.vijk[3 * u + 1]) +
                   abs(b.vijk[3 * v + 2] - b.vijk[3 * u + 2]);
    if (li > sc.colNearHops && !(sc.flags & SC_ATTACH)) return true;   // attached links shortcut the lattice
    return !nearbyOf(b, sc, v, u);
}

// Hands each active lane a range of n entries from its simulation's counter (*counter of
// sim s), one atomic per simulation present in the wave; returns the range's start (an
// offset in the simulation's table).  Ranges are contiguous per lane; their order across
// voxels is arbitrary (only each voxel's own list order matters).
template <typename CounterOf>
__device__ __forceinline__ int waveAlloc(bool active, int s, int n, CounterOf counter) {
    const int lane = threadIdx.x & 63;
    int start = 0;
    unsigned long long pending = __ballot(active);
    while (pending) {
        const int lead = __ffsll((unsigned long long)pending) - 1;
        const int sl = __shfl(s, lead, 64);
        const bool mine = active && s == sl;
        const int x = mine ? n : 0;
        int incl = x;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        const int tot = __shfl(incl, 63, 64);
        int base = 0;
        if (lane == lead) base = atomicAdd(counter(sl), tot);
        base = __shfl(base, lead, 64);
        if (mine) start = base + incl - x;
        pending &= ~__ballot(mine);
    }
    return start;
}

// ---- nearby lists after a topology change (CVX_Voxel::generateNearby, Voxelyze.cpp:394-417,
// on the links present now).  A step that attaches, detaches or removes voxels marks the
// simulation (SC_TOPO_CHG); the next step regenerates the nearby list of every surface
// voxel before its collision rebuild (the CPU engine's nearbyStale -> collisionsStale),
// after its link pass, so that the links that pass detached are already gone, as in
// CVoxelyze::updateCollisions.  One thread per voxel runs the breadth-first search of
// generateNearby with its queue in LDS (VX3_NEAR_MAX voxels); count, scan and emit
// passes fill nearList.  A search or a table too small for the result stops the
// simulation after its link pass (SC_COL_GROW | SC_LINKS_DONE) until the host builds
// the lists itself or enlarges the table.
#define VX3_NEAR_MAX 384
#define VX3_NEAR_BS 64

__device__ __forceinline__ bool nearDirty(const DevBatch<real>& b, int s) {
    return simRunning(b, s) && (b.sims[s].flags & SC_NEARDIRTY) && !(b.sims[s].flags & SC_DIVERGED_PENDING);
}

// Queues v and the voxels within `hops` links of it, level by level, into q; returns
// how many, or -1 when they do not fit.
__device__ int nearBfs(const DevBatch<real>& b, int v, int hops, int* q) {
    const size_t V = (size_t)b.Vcap;
    int n = 0, cur = 0;
    q[n++] = v;
    for (int k = 0; k < hops; k++) {
        const int end = n;
        while (cur < end) {
            const int u = q[cur++];
            const int fl = b.vflags[u];
            for (int d = 0; d < 6; d++) {
                if (!(fl & (1 << (VF_LINKMASK_SHIFT + d)))) continue;
                const int w = b.nbr[d * V + u];
                bool seen = false;
                for (int i = 0; i < n && !seen; i++) seen = q[i] == w;
                if (seen) continue;
                if (n == VX3_NEAR_MAX) return -1;
                q[n++] = w;
            }
        }
    }
    return n;
}

__global__ void __launch_bounds__(VX3_NEAR_BS) k_near_count(DevBatch<real> b) {
    __shared__ int q[VX3_NEAR_BS][VX3_NEAR_MAX];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = v < b.vHi ? b.vsim[v] : -1;
    const bool act = v < b.vHi && nearDirty(b, s) && !(b.vflags[v] & VF_INTERIOR);
    int c = 0;
    bool over = false;
    if (act) {
        int* Q = q[threadIdx.x];
        const int n = nearBfs(b, v, b.sims[s].colNearHops, Q);
        if (n < 0) over = true;
        else for (int i = 1; i < n; i++) c += !(b.vflags[Q[i]] & VF_INTERIOR);
    }
    const int start = waveAlloc(act, s, c, [&](int sl) { return &b.sims[sl].nearUsed; });
    if (v >= b.vHi || !nearDirty(b, s)) return;
    b.nearCount[v] = act ? c : 0;
    b.nearStart[v] = b.sims[s].near0 + start;
    if (over) {
        atomicMax(&b.sims[s].nearNeed, 1 << 30);   // the host builds this simulation's lists itself
        atomicOr(&b.sims[s].flags, SC_COL_GROW | SC_LINKS_DONE);
    } else if (start + c > b.sims[s].nearCap) {
        atomicMax(&b.sims[s].nearNeed, start + c);
        atomicOr(&b.sims[s].flags, SC_COL_GROW | SC_LINKS_DONE);
    }
}

__global__ void __launch_bounds__(VX3_NEAR_BS) k_near_emit(DevBatch<real> b) {
    __shared__ int q[VX3_NEAR_BS][VX3_NEAR_MAX];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= b.vHi) return;
    const int s = b.vsim[v];
    if (!nearDirty(b, s) || (b.sims[s].flags & SC_COL_GROW)) return;
    if (b.vflags[v] & VF_INTERIOR) {
        b.nearCount[v] = 0;
        if (b.sims[s].flags & SC_NEARMASK)
            for (int w = 0; w < VX3_NEAR_WORDS; w++) b.nearMask[(size_t)w * b.Vcap + v] = 0;
        return;
    }
    int* Q = q[threadIdx.x];
    const int n = nearBfs(b, v, b.sims[s].colNearHops, Q);
    int* out = b.nearList + b.nearStart[v];
    const int v0 = b.sims[s].v0;
    const bool mask = (b.sims[s].flags & SC_NEARMASK) != 0;
    unsigned mk[VX3_NEAR_WORDS] = {0, 0, 0, 0, 0, 0, 0, 0};
    int c = 0;
    for (int i = 1; i < n; i++)
        if (!(b.vflags[Q[i]] & VF_INTERIOR)) {
            out[c++] = Q[i] - v0;
            if (mask) {   // within the hop radius (<= 5), so within the table's offsets
                const int u = Q[i];
                const int bit = kNearBits.bit[(b.vijk[3 * u] - b.vijk[3 * v] + 5) * 121 + (b.vijk[3 * u + 1] - b.vijk[3 * v + 1] + 5) * 11 +
                                              (b.vijk[3 * u + 2] - b.vijk[3 * v + 2] + 5)];
                mk[bit >> 5] |= 1u << (bit & 31);
            }
        }
    if (mask)
        for (int w = 0; w < VX3_NEAR_WORDS; w++) b.nearMask[(size_t)w * b.Vcap + v] = mk[w];
    for (int i = 1; i < c; i++) {   // ascending, for the binary search of colWatched
        const int x = out[i];
        int j = i - 1;
        while (j >= 0 && out[j] > x) { out[j + 1] = out[j]; j--; }
        out[j + 1] = x;
    }
}

// Takes the contact snapshot of every surface voxel and decides, per simulation,
// whether its lists are rebuilt this step: the first step, every colInterval steps,
// after new nearby lists, or once a surface voxel moved more than recalcDist since the
// last rebuild
__device__ __forceinline__ void colStalePass(const DevBatch<real>& b, int v) {
    int s = -1;
    bool need = false, rebuilding = false;
    if (v < b.vHi && colVoxel(b, v, s)) {
        const SimCtl& sc = b.sims[s];
        // contact snapshot (see DevBatch::cpx): the reference computes every contact force
        // from the state before any voxel moves this step (CVoxelyze::updateCollisions ahead
        // of the voxel loop, Voxelyze.cpp:263-279); the link pass before this one moves none
        b.cpx[v] = b.px[v]; b.cpy[v] = b.py[v]; b.cpz[v] = b.pz[v];
        b.cmx[v] = b.mx[v]; b.cmy[v] = b.my[v]; b.cmz[v] = b.mz[v];
        b.ctf[v] = b.tfac[v];
        rebuilding = (sc.flags & SC_REBUILD) != 0;
        if (!rebuilding) {
            // new nearby lists (CVoxelyze: nearbyStale -> collisionsStale), or the fixed period
            // (both normally set by the control pass of the step before)
            need = (sc.flags & SC_NEARDIRTY) || (sc.colInterval > 0 && sc.steps % sc.colInterval == 0);
            if (!need) {
                const real dx = b.px[v] - (real)b.lastWatch[3 * v], dy = b.py[v] - (real)b.lastWatch[3 * v + 1],
                           dz = b.pz[v] - (real)b.lastWatch[3 * v + 2];
                need = dx * dx + dy * dy + dz * dz > (real)sc.colRecalcSq;
            }
        }
    }
    // the rest of the rebuild passes run only if some simulation rebuilds this step
    if (__ballot(need || rebuilding) && (threadIdx.x & 63) == 0) *b.colAny = 1;
    // one atomic per simulation present in the wave, not one per voxel
    unsigned long long pending = __ballot(need);
    while (pending) {
        const int lead = __ffsll((unsigned long long)pending) - 1;
        const int sl = __shfl(s, lead, 64);
        if ((int)(threadIdx.x & 63) == lead) atomicOr(&b.sims[sl].flags, SC_REBUILD);
        pending &= ~__ballot(need && s == sl);
    }
}

// The staleness pass and the link pass in one launch (one thread per voxel and per link):
// they are independent -- the link pass moves no voxel and changes no thermal factor,
// and the staleness pass reads only voxel state and the control blocks' collision
// bits -- and an idle launch costs about as much as the staleness pass itself.  The
// same launch zeroes every bucket count: only a rebuild reads them, and a rebuild
// counts from zero in this step's later passes.
__global__ void __launch_bounds__(256) k_links_stale(DevBatch<real> b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    colStalePass(b, i);   // every lane of the wave: it votes
    for (int h = i; h < b.hashTot; h += gridDim.x * blockDim.x) b.hashCnt[h] = 0;
    linkPass(b, i);
}

__device__ __forceinline__ bool colRebuilding(const DevBatch<real>& b, int s) {
    if (s < 0 || !simRunning(b, s)) return false;
    const int f = b.sims[s].flags;
    return (f & SC_REBUILD) && (f & SC_COLLISIONS) && !(f & (SC_DIVERGED_PENDING | SC_COL_GROW));
}

__global__ void __launch_bounds__(256) k_col_insert(DevBatch<real> b) {
    if (!*b.colAny) return;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= b.vHi) return;
    int s;
    if (!colVoxel(b, v, s)) return;
    SimCtl& sc = b.sims[s];
    if (!(sc.flags & SC_REBUILD)) return;
    b.lastWatch[3 * v] = (float)b.px[v];
    b.lastWatch[3 * v + 1] = (float)b.py[v];
    b.lastWatch[3 * v + 2] = (float)b.pz[v];
    int c[3];
    colCell(b, v, sc.colCell, c);
    const int h = sc.h0 + (int)(colHash(c[0], c[1], c[2]) & ((1u << sc.hBits) - 1));
    b.vbk[v] = h;
    b.vrank[v] = atomicAdd(&b.hashCnt[h], 1);
}

// exclusive scan of the bucket counts within each block of VX3_COL_HBLOCK buckets
// (a simulation's buckets are whole blocks); the block's total goes to hashBlk
__global__ void __launch_bounds__(256) k_col_bscan(DevBatch<real> b) {
    if (!*b.colAny) return;
    const int B = blockIdx.x;
    const int h0 = B * VX3_COL_HBLOCK;
    if (h0 >= b.hashTot || !colRebuilding(b, b.hashSim[h0])) return;
    constexpr int PER = VX3_COL_HBLOCK / 256;
    __shared__ int part[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int x[PER], sum = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) { x[k] = b.hashCnt[h0 + tid * PER + k]; sum += x[k]; }
    int incl = sum;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) part[w] = incl;
    __syncthreads();
    int before = incl - sum;
    for (int i = 0; i < w; i++) before += part[i];
#pragma unroll
    for (int k = 0; k < PER; k++) { b.hashStart[h0 + tid * PER + k] = before; before += x[k]; }
    // The block's total, then a count of the simulation's scanned blocks; the block that
    // counts last turns the totals into offsets.  The total is an agent-scope store that
    // has completed (s_waitcnt 0) before the count, and the last block reads the totals
    // with agent-scope loads, so no cache writeback is needed between them.
    if (tid == 255) {
        __hip_atomic_store(&b.hashBlk[B], before, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);
    }
    const int s = b.hashSim[h0];
    const SimCtl& sc = b.sims[s];
    const int nB = (1 << sc.hBits) / VX3_COL_HBLOCK;
    __shared__ int last;
    __syncthreads();
    if (tid == 0) last = __hip_atomic_fetch_add(&b.sims[s].scanDone, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nB - 1;
    __syncthreads();
    if (!last) return;
    // per rebuilt simulation: offsets of its bucket blocks in its cellVox range
    __shared__ int carry;
    const int B0 = sc.h0 / VX3_COL_HBLOCK;
    if (tid == 0) {
        carry = sc.c0;
        b.sims[s].scanDone = 0;   // for the next rebuild (a later launch)
        b.sims[s].nbUsed = 0;     // the partner ranges of this rebuild are handed out from 0
        b.sims[s].rebuilds++;
    }
    __syncthreads();
    for (int base = 0; base < nB; base += 256) {
        const int i = base + tid;
        const int tot = i < nB ? __hip_atomic_load(&b.hashBlk[B0 + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
        int run = tot;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(run, o, 64);
            if (lane >= o) run += y;
        }
        if (lane == 63) part[w] = run;
        __syncthreads();
        int off = carry + run - tot;
        for (int q = 0; q < w; q++) off += part[q];
        if (i < nB) b.hashBlkOff[B0 + i] = off;
        __syncthreads();
        if (tid == 255) carry = off + tot;
        __syncthreads();
    }
}

__device__ __forceinline__ int colBucketStart(const DevBatch<real>& b, int h) {
    return b.hashBlkOff[h / VX3_COL_HBLOCK] + b.hashStart[h];
}

__global__ void __launch_bounds__(256) k_col_place(DevBatch<real> b) {
    if (!*b.colAny) return;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= b.vHi) return;
    int s;
    if (!colVoxel(b, v, s) || !(b.sims[s].flags & SC_REBUILD)) return;
    const int h = b.vbk[v];
    b.cellVox[colBucketStart(b, h) + b.vrank[v]] = v;
}

// visits every watched partner u of surface voxel v once (any order)
//
// Before the exact test, a box test on the float positions of this rebuild (lastWatch:
// k_col_insert writes them for every voxel it puts in the buckets, earlier in the same
// step), one 12-byte read per candidate instead of three reals from three arrays.  It
// only drops candidates further than the watch radius along one axis by more than float
// rounding can explain: with a = the voxel's coordinate, b = the candidate's and both
// rounded to float, the float difference is within 2^-24 (|a| + |b| + |difference|) of
// a - b, which for a watched pair (|a - b| <= r) is below 2^-22 (|a| + r); the bound
// used is twice that.  Candidates are read four at a time so their loads overlap.
template <typename F>
__device__ __forceinline__ void colForPartners(const DevBatch<real>& b, const SimCtl& sc, int v, F f) {
    int c[3];
    colCell(b, v, sc.colCell, c);
    const unsigned mask = (1u << sc.hBits) - 1;
    const float* lw = b.lastWatch;
    const float vx = lw[3 * v], vy = lw[3 * v + 1], vz = lw[3 * v + 2];
    const float r = sqrtf(sc.colThreshSq) * (1.0f + 0x1p-20f);
    const float rx = r + (fabsf(vx) + r) * 0x1p-21f, ry = r + (fabsf(vy) + r) * 0x1p-21f, rz = r + (fabsf(vz) + r) * 0x1p-21f;
    int hs[27];
#pragma unroll
    for (int k = 0; k < 27; k++)
        hs[k] = sc.h0 + (int)(colHash(c[0] + k % 3 - 1, c[1] + (k / 3) % 3 - 1, c[2] + k / 9 - 1) & mask);
    for (int k = 0; k < 27; k++) {
        const int h = hs[k];
        bool seen = false;   // a bucket shared by several of the 27 cells is read once
        for (int j = 0; j < k; j++) seen |= hs[j] == h;
        if (seen) continue;
        const int n = b.hashCnt[h];
        if (!n) continue;
        const int* cv = b.cellVox + colBucketStart(b, h);
        for (int i0 = 0; i0 < n; i0 += 4) {
            int u[4];
            float px[4], py[4], pz[4];
#pragma unroll
            for (int q = 0; q < 4; q++) u[q] = i0 + q < n ? cv[i0 + q] : -1;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int w = u[q] < 0 ? v : u[q];
                px[q] = lw[3 * w]; py[q] = lw[3 * w + 1]; pz[q] = lw[3 * w + 2];
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (u[q] >= 0 && u[q] != v && fabsf(vx - px[q]) <= rx && fabsf(vy - py[q]) <= ry && fabsf(vz - pz[q]) <= rz &&
                    colWatched(b, sc, v, u[q]))
                    f(u[q]);
        }
    }
}

// Partner lists of every surface voxel being rebuilt, in one pass over the candidate
// cells: the partners are gathered into an LDS column of the voxel's thread (up to
// VX3_COL_STAGE; beyond that the cells are walked a second time straight into the table),
// the voxel receives a range of the neighbour table, and the list is sorted there.  A
// table too small stops the simulation after its link pass until the host grows it.
// 32 entries (32 KB per workgroup) keep five workgroups on a CU, the register limit
#define VX3_COL_STAGE 32
//
// One thread per cellVox entry rather than per voxel: every lane then holds a surface
// voxel (interior voxels have no list), and lanes of one cell share their 27 buckets.
// An entry is live when k_col_place put its voxel there in this step's rebuild; entries
// past a simulation's current count keep voxels of an older rebuild (or -1).
__global__ void __launch_bounds__(256) k_col_list(DevBatch<real> b) {
    if (!*b.colAny) return;
    __shared__ int stage[VX3_COL_STAGE][256];   // [entry][thread]: a thread's column, conflict-free
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = threadIdx.x;
    int s = -1;
    const int v = e < b.cellHi ? b.cellVox[e] : -1;
    const bool act = v >= 0 && v < b.vHi && colVoxel(b, v, s) && (b.sims[s].flags & SC_REBUILD) &&
                     colBucketStart(b, b.vbk[v]) + b.vrank[v] == e;
    int n = 0;
    if (act)
        colForPartners(b, b.sims[s], v, [&](int u) {
            if (n < VX3_COL_STAGE) stage[n][t] = u;
            n++;
        });
    const int start = waveAlloc(act, s, n, [&](int sl) { return &b.sims[sl].nbUsed; });
    if (!act) return;
    SimCtl& sc = b.sims[s];
    b.nbCount[v] = n;
    b.nbStart[v] = sc.nb0 + start;
    if (start + n > sc.nbCap) {
        atomicMax(&sc.nbNeed, start + n);
        atomicOr(&sc.flags, SC_COL_GROW | SC_LINKS_DONE);
        return;
    }
    int* out = b.nbList + sc.nb0 + start;
    if (n <= VX3_COL_STAGE) {
        for (int i = 1; i < n; i++) {   // ascending partner order
            const int x = stage[i][t];
            int j = i - 1;
            while (j >= 0 && stage[j][t] > x) { stage[j + 1][t] = stage[j][t]; j--; }
            stage[j + 1][t] = x;
        }
        for (int i = 0; i < n; i++) out[i] = stage[i][t];
    } else {
        int m = 0;
        colForPartners(b, sc, v, [&](int u) { out[m++] = u; });
        for (int i = 1; i < n; i++) {
            const int x = out[i];
            int j = i - 1;
            while (j >= 0 && out[j] > x) { out[j + 1] = out[j]; j--; }
            out[j + 1] = x;
        }
    }
}


// Contact force of the pair (i, j), i the lower index, on voxel i: CVX_Collision::
// updateContactForce (VX_Collision.cpp:41-56) from both voxels' pre-step position,
// momentum and thermal factor.  Shared by the per-step and resident kernels.
__device__ __forceinline__ void pairContact(const VoxMatDev& m1, const VoxMatDev& m2, real pix, real piy, real piz, real pjx,
                                            real pjy, real pjz, real mix, real miy, real miz, real mjx, real mjy, real mjz,
                                            float tfi, float tfj, real nom, float& fx, float& fy, float& fz) {
    const float kp = 2.0f / (1.0f / m1.penetrationStiffness + 1.0f / m2.penetrationStiffness);
    const float cd = 0.5f * (m1.zetaCollision * m1.c2xSqMxExS + m2.zetaCollision * m2.c2xSqMxExS);
    const float ox = (float)(pjx - pix), oy = (float)(pjy - piy), oz = (float)(pjz - piz);
    const real b1 = nom * (real)tfi, b2 = nom * (real)tfj;
    const float nomDist = (float)(((b1 + b1 + b1) / (real)3.0f + (b2 + b2 + b2) / (real)3.0f) * (real)0.625f);
    const float len = sqrtf(ox * ox + oy * oy + oz * oz);
    const float rel = nomDist - len;
    fx = fy = fz = 0;
    if (rel > 0) {
        float ux = ox, uy = oy, uz = oz;
        if (len > 0) { const float inv = 1.0f / len; ux = inv * ox; uy = inv * oy; uz = inv * oz; }
        const real mi1 = (real)m1.massInverse, mi2 = (real)m2.massInverse;
        const real v1 = (mix * mi1) * (real)ux + (miy * mi1) * (real)uy + (miz * mi1) * (real)uz;
        const real v2 = (mjx * mi2) * (real)ux + (mjy * mi2) * (real)uy + (mjz * mi2) * (real)uz;
        const float relVel = (float)(v1 - v2);
        const float mag = kp * rel + cd * relVel;
        fx = ux * mag; fy = uy * mag; fz = uz * mag;
    }
}

// contact force on voxel a from its pair with c, from the pair's first voxel
// min(a, c): CVX_Collision::updateContactForce (VX_Collision.cpp:41-56) and
// contactForce(pVoxel) (VX_Collision.h:44), subtracted in the voxel's partner order
struct Contacts {
    const DevBatch<real>* b;
    int v, start, count;
    real nom;
    template <typename T> __device__ __forceinline__ void subtract(V3<T>& force) const {
        for (int k = 0; k < count; k++) {
            const int u = b->nbList[start + k];
            const int i = v < u ? v : u, j = v < u ? u : v;
            float fx, fy, fz;
            // both voxels' state from the contact snapshot (pre-step values)
            pairContact(b->vmatTab[b->vmat[i]], b->vmatTab[b->vmat[j]], b->cpx[i], b->cpy[i], b->cpz[i], b->cpx[j], b->cpy[j],
                        b->cpz[j], b->cmx[i], b->cmy[i], b->cmz[i], b->cmx[j], b->cmy[j], b->cmz[j], b->ctf[i], b->ctf[j], nom,
                        fx, fy, fz);
            if (v != i) { fx = -fx; fy = -fy; fz = -fz; }
            force.x -= (T)fx; force.y -= (T)fy; force.z -= (T)fz;
        }
    }
};

// ---- signals (VX3_Voxel.cu:279-348).  sg[k*V + v]: 0 local signal, 1 its next decay
// time, 2 next pacemaker pulse, 3 deaf until, 4/5 pending signal value/time, 6/7 posted
// value/time; sp[k*V + v]: 0 pacemaker, 1 period, 2 value decay, 3 time delay, 4 inactive period.
// VX3_Voxel::receiveSignal
__device__ __forceinline__ void receiveSignal(const DevBatch<real>& b, size_t V, int v, double value, double activeTime, bool force) {
    if (!force && b.sg[3 * V + v] > activeTime) return;
    if (value < 0.1) return;
    b.sg[3 * V + v] = activeTime + b.sp[4 * V + v];
    b.sg[0 * V + v] = value;
    double pv = value * b.sp[2 * V + v];
    if (pv < 0.1) pv = 0;
    b.sg[4 * V + v] = pv;
    b.sg[5 * V + v] = activeTime;
}
// End of a moving voxel's update: propagateSignal (its pending signal is posted for the
// neighbours), packMaker, localSignalDecay -- all on the voxel's own state
__device__ __forceinline__ void signalsOwn(const DevBatch<real>& b, size_t V, int v, double t) {
    double outV = 0, outT = 0;
    if (!(b.sg[5 * V + v] > t) && !(b.sg[4 * V + v] < 0.1)) {
        outV = b.sg[4 * V + v];
        outT = t + b.sp[3 * V + v];
        b.sg[4 * V + v] = 0;
        b.sg[5 * V + v] = 0;
        b.sg[3 * V + v] = t + 2 * b.sp[3 * V + v] + b.sp[4 * V + v];
    }
    if (b.sp[0 * V + v] != 0 && !(b.sg[2 * V + v] > t)) {
        receiveSignal(b, V, v, 100, t, true);
        b.sg[2 * V + v] = t + b.sp[1 * V + v];
    }
    if (!(b.sg[1 * V + v] > t)) {
        if (b.sg[0 * V + v] < 0.1) b.sg[0 * V + v] = 0;
        else {
            b.sg[0 * V + v] = b.sg[0 * V + v] * 0.9;
            b.sg[1 * V + v] = t + 0.01;
        }
    }
    b.sg[6 * V + v] = outV;
    b.sg[7 * V + v] = outT;
}

// The extra forces of a simulation with cilia or a force field, added where the
// reference adds them: the cilia force at the end of force() (VX3_Voxel.cu:376), the
// force field right after force() returns (VX3_Voxel.cu:178-184).
template <typename Inner> struct WithExtras {
    Inner inner;
    const real* xf;
    size_t V;
    int v;
    template <typename T> __device__ __forceinline__ void subtract(V3<T>& f) const {
        inner.subtract(f);
        f.add(V3<T>(xf[0 * V + v], xf[1 * V + v], xf[2 * V + v]));
        f.x += xf[3 * V + v];
        f.y += xf[4 * V + v];
        f.z += xf[5 * V + v];
    }
};

// The voxel pass of the simulations without (EXTRAS false) or with (true) cilia or a
// force field; the two are separate kernels so that the common one keeps its
// registers.
template <bool EXTRAS> __device__ __forceinline__ void voxelPass(const DevBatch<real>& b) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= b.vHi) return;
    const int s = b.vsim[v];
    if (!simRunning(b, s)) return;
    const SimCtl& sc = b.sims[s];
    if (((sc.flags & SC_EXTRAS) != 0) != EXTRAS) return;
    if (sc.flags & (SC_DIVERGED_PENDING | SC_COL_OVERFLOW)) return;   // doTimeStep returns before the voxel loop
    if ((sc.flags & SC_SECONDARY) && (b.vflags[v] & VF_REMOVED)) return;   // VX3_VoxelyzeKernel.cu:586,632
    const VoxMatDev& m = b.vmatTab[b.vmat[v]];
    const float tf = b.tfac[v];
    if (!m.fixed) {
        int fl = b.vflags[v];
        V3<real> f(0, 0, 0), mo(0, 0, 0);
        const size_t V = (size_t)b.Vcap;
        const real* F = b.F;
        const size_t FS = b.Fs;
#pragma unroll
        for (int d = 0; d < 6; d++) {
            if (fl & (1 << (VF_LINKMASK_SHIFT + d))) {
                f.add(V3<real>(F[(d * 6 + 0) * FS + v], F[(d * 6 + 1) * FS + v], F[(d * 6 + 2) * FS + v]));
                mo.add(V3<real>(F[(d * 6 + 3) * FS + v], F[(d * 6 + 4) * FS + v], F[(d * 6 + 5) * FS + v]));
            }
        }
        const int e = b.vext[v];
        const ExtDev* ext = e >= 0 ? &b.extTab[e] : nullptr;
        int ix = 0, iy = 0, iz = 0;
        if (ext) { ix = ext->ix; iy = ext->iy; iz = ext->iz; }
        V3<real> pos(b.px[v], b.py[v], b.pz[v]), lm(b.mx[v], b.my[v], b.mz[v]), am(b.ax[v], b.ay[v], b.az[v]);
        Q4<real> o(b.ow[v], b.ox[v], b.oy[v], b.oz[v]);
        if constexpr (EXTRAS) {
            if ((sc.flags & SC_COLLISIONS) && !(fl & VF_INTERIOR)) {
                WithExtras<Contacts> con{Contacts{&b, v, b.nbStart[v], b.nbCount[v], (real)sc.voxSize}, b.xf, V, v};
                voxelStep<real>(m, ext, sc.dt, (real)sc.voxSize, ix, iy, iz, tf, f, mo, con, pos, lm, o, am, fl);
            } else {
                WithExtras<NoContacts> con{NoContacts(), b.xf, V, v};
                voxelStep<real>(m, ext, sc.dt, (real)sc.voxSize, ix, iy, iz, tf, f, mo, con, pos, lm, o, am, fl);
            }
        } else {
            if ((sc.flags & SC_COLLISIONS) && !(fl & VF_INTERIOR)) {
                Contacts con{&b, v, b.nbStart[v], b.nbCount[v], (real)sc.voxSize};
                voxelStep<real>(m, ext, sc.dt, (real)sc.voxSize, ix, iy, iz, tf, f, mo, con, pos, lm, o, am, fl);
            } else {
                voxelStep<real>(m, ext, sc.dt, (real)sc.voxSize, ix, iy, iz, tf, f, mo, NoContacts(), pos, lm, o, am, fl);
            }
        }
        b.px[v] = pos.x; b.py[v] = pos.y; b.pz[v] = pos.z;
        b.mx[v] = lm.x; b.my[v] = lm.y; b.mz[v] = lm.z;
        b.ow[v] = o.w; b.ox[v] = o.x; b.oy[v] = o.y; b.oz[v] = o.z;
        b.ax[v] = am.x; b.ay[v] = am.y; b.az[v] = am.z;
        b.vflags[v] = fl;
        if constexpr (EXTRAS) {
            if (sc.flags & SC_SIGNALS) signalsOwn(b, V, v, sc.t);
        }
    }
    // thermal factor for the next step: VX3 updateTemperature at t + dt (VX3_VoxelyzeKernel.cu:625-650)
    if ((sc.flags & SC_VARY) && !m.fixed) {
        const double tn = sc.t + (double)sc.dt;
        if (!(m.thermalOnAfter > tn)) {
            const float temp = actuationTemp(sc.tempAmp, sc.tempPeriod, b.phase[v], tn, (sc.flags & SC_EXPANSION) != 0);
            b.tfac[v] = 1 + temp * m.alphaCTE;
        }
    }
    if (sc.flags & SC_POISSON) poissonStrain(b, v);
}

// the 36 end-load slots of voxels [0, nv) from one buffer to another (their strides may differ)
__global__ void __launch_bounds__(256) k_copy_loads(const real* __restrict__ src, size_t srcFs, real* __restrict__ dst,
                                                    size_t dstFs, int nv) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
#pragma unroll
    for (int k = 0; k < 36; k++) dst[k * dstFs + v] = src[k * srcFs + v];
}

// p[i] = x for i in [0, n) step `stride` (a kernel, not a memset: callers may capture it in a graph)
__global__ void __launch_bounds__(256) k_fill_ints(int* p, int n, int stride, int x) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) * stride;
    if (i < n) p[i] = x;
}

// at most 128 VGPRs (four waves per SIMD): the voxel pass is a gather over the force
// slots, latency-bound at the three waves its unconstrained allocation allows
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_voxels(DevBatch<real> b) { voxelPass<false>(b); }
__global__ void __launch_bounds__(256) k_voxels_extras(DevBatch<real> b) { voxelPass<true>(b); }

// ---- MathTree (queue evaluator of Utils/VX3_MathTree.h:50-192), device side, bounded programs.
// The value queue q lives in LDS (dynamically indexed), supplied by the caller.
#define VX3_DEV_MT_MAX 64
// Transcendental operators out of line, so that the evaluator inlined into the step
// kernels stays small in registers.
__device__ __noinline__ double mtFunc(int op, double a, double b) {
    switch (op) {
    case 9: return pow(a, b);
    case 10: return sqrt(b);
    case 11: return sin(b);
    case 12: return cos(b);
    case 13: return tan(b);
    case 14: return atan(b);
    case 15: return log(b);
    case 23: return 0.5 * erfc(-b / sqrt(2.0));
    default: return 0.0;
    }
}
__device__ __forceinline__ double mathTreeEval(const TokenDev* buf, int n, double x, double y, double z, double hit, double t,
                                               double angle, double closeness, int ncp, int nvox, double* q) {
    if (n <= 0) return 0.0;
    int vc = 0, pc = 0;
    for (int i = 0; i < n && i < VX3_DEV_MT_MAX; i++) {
        const int op = buf[i].op;
        const double val = buf[i].value;
        double r = 0;
        switch (op) {
        case 0: return q[pc];
        case 1: r = val; break;
        case 2: r = 2.71828182845904523536; break;
        case 3: r = 3.14159265358979323846; break;
        case 4: {
            int k = val < 0.5 ? 0 : (int)(val + 0.5);
            r = k == 0 ? x : k == 1 ? y : k == 2 ? z : k == 3 ? hit : k == 4 ? t : k == 5 ? angle : k == 6 ? closeness
                : k == 7 ? (double)ncp : k == 8 ? (double)nvox : 0.0;
            break;
        }
        case 5: r = q[pc + 1] + q[pc]; pc += 2; break;
        case 6: r = q[pc + 1] - q[pc]; pc += 2; break;
        case 7: r = q[pc + 1] * q[pc]; pc += 2; break;
        case 8: r = q[pc + 1] / q[pc]; pc += 2; break;
        case 9: r = mtFunc(9, q[pc + 1], q[pc]); pc += 2; break;
        case 10: r = mtFunc(10, 0.0, q[pc]); pc++; break;
        case 11: r = mtFunc(11, 0.0, q[pc]); pc++; break;
        case 12: r = mtFunc(12, 0.0, q[pc]); pc++; break;
        case 13: r = mtFunc(13, 0.0, q[pc]); pc++; break;
        case 14: r = mtFunc(14, 0.0, q[pc]); pc++; break;
        case 15: r = mtFunc(15, 0.0, q[pc]); pc++; break;
        case 16: r = rint(q[pc]); pc++; break;
        case 17: r = fabs(q[pc]); pc++; break;
        case 18: r = !q[pc]; pc++; break;
        case 19: r = q[pc + 1] > q[pc]; pc += 2; break;
        case 20: r = q[pc + 1] < q[pc]; pc += 2; break;
        case 21: r = q[pc + 1] && q[pc]; pc += 2; break;
        case 22: r = q[pc + 1] || q[pc]; pc += 2; break;
        case 23: r = mtFunc(23, 0.0, q[pc]); pc++; break;
        default: return -1;
        }
        if (pc > vc) return -1;
        q[vc++] = r;
    }
    return q[pc < VX3_DEV_MT_MAX ? pc : 0];
}

// Step passes of the simulations with VX3 detach, removal, cilia or a force field.
// forces == 0, at the step's start: a voxel's surface state follows its link mask as
// of the previous step's end (regenerateSurfaceVoxels, VX3_VoxelyzeKernel.cu:281-285).
// forces == 1, before the voxel pass: extra global forces, from the state before this
// step's update:
// cilia as gpu_update_cilia_force (VX3_VoxelyzeKernel.cu:846-859): a surface voxel
// whose material has Cilia != 0 and whose switch-on time has passed is pushed by
// orient.RotateVec3D(base + localSignal*shift) * Cilia, localSignal being 0 (signals
// are not simulated); the force field's three programs of the voxel position at the
// step's start time (VX3_Voxel.cu:178-184; hit, closeness and close pairs as of the last
// sample).
__global__ void __launch_bounds__(256) k_prestep(DevBatch<real> b, int forces) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= b.vHi) return;
    const int s = b.vsim[v];
    if (!simRunning(b, s)) return;
    const SimCtl& sc = b.sims[s];
    if (!forces) {   // before the link pass: CVX_Voxel::updateSurface, interior iff all six links remain
        if (!(sc.flags & (SC_DETACH | SC_SECONDARY | SC_ATTACH)) || (sc.flags & SC_LINKS_DONE)) return;
        const int fl = b.vflags[v];
        const int all = 0x3F << VF_LINKMASK_SHIFT;
        const int nf = (fl & all) == all ? (fl | VF_INTERIOR) : (fl & ~VF_INTERIOR);
        if (nf != fl) b.vflags[v] = nf;
        return;
    }
    if (forces == 2) {   // after the voxel pass: VX3_Voxel::enableAttach (VX3_VoxelyzeKernel.cu:608-621)
        if (!(sc.flags & SC_ATTACH) || (sc.flags & (SC_DIVERGED_PENDING | SC_COL_OVERFLOW))) return;
        const int fl = b.vflags[v];
        if ((fl & VF_REMOVED) || b.vmatTab[b.vmat[v]].fixed) return;   // not updated, keeps its value
        bool on = true;
        double q[VX3_DEV_MT_MAX + 1];
        for (int k = 0; k < 5 && on; k++)   // all five conditions > 0; a missing one does not restrict
            if (sc.atLen[k] > 0)
                on = mathTreeEval(b.tokens + sc.atOff[k], sc.atLen[k], (double)b.px[v], (double)b.py[v], (double)b.pz[v],
                                  (double)sc.hits, sc.t, sc.recentAngle, sc.closeness, sc.closePairs, sc.nv, q) > 0;
        const int nf = on ? (fl | VF_ATTACH) : (fl & ~VF_ATTACH);
        if (nf != fl) b.vflags[v] = nf;
        return;
    }
    // after the link and contact passes (gpu_update_cilia_force runs after updateAttach)
    if (!(sc.flags & SC_EXTRAS) || (sc.flags & (SC_DIVERGED_PENDING | SC_COL_OVERFLOW))) return;
    const size_t V = (size_t)b.Vcap;
    V3<real> c(0, 0, 0);
    if (sc.flags & SC_CILIA) {
        const double coef = b.cil[6 * V + v], on = b.cil[7 * V + v];
        if (!(b.vflags[v] & (VF_INTERIOR | VF_REMOVED)) && coef != 0 && !(on > sc.t)) {
            const double sig = (sc.flags & SC_SIGNALS) ? b.sg[v] : 0.0;   // localSignal
            const V3<real> f((real)(b.cil[0 * V + v] + sig * b.cil[3 * V + v]), (real)(b.cil[1 * V + v] + sig * b.cil[4 * V + v]),
                             (real)(b.cil[2 * V + v] + sig * b.cil[5 * V + v]));
            const Q4<real> o(b.ow[v], b.ox[v], b.oy[v], b.oz[v]);
            c = o.rot(f).mul((real)coef);
        }
    }
    double ff[3] = {0, 0, 0};
    double q[VX3_DEV_MT_MAX + 1];
    for (int k = 0; k < 3; k++)
        if (sc.ffLen[k] > 0)
            ff[k] = mathTreeEval(b.tokens + sc.ffOff[k], sc.ffLen[k], (double)b.px[v], (double)b.py[v], (double)b.pz[v],
                                 (double)sc.hits, sc.t, sc.recentAngle, sc.closeness, sc.closePairs, sc.nv, q);
    b.xf[0 * V + v] = c.x; b.xf[1 * V + v] = c.y; b.xf[2 * V + v] = c.z;
    b.xf[3 * V + v] = (real)ff[0]; b.xf[4 * V + v] = (real)ff[1]; b.xf[5 * V + v] = (real)ff[2];
}

// ---- VX3 attach (handle_collision_attachment, VX3_VoxelyzeKernel.cu:724-826, over the
// surface pairs of gpu_update_attach, :833-843)
//
// One workgroup per simulation.  Only voxels that can attach take part: on the surface,
// not removed, their attach conditions held at their last update, of a sticky material
// that is not fixed.  Those are compacted in voxel order into the simulation's attCand
// range; then the candidate pairs (higher, lower) -- within the envelope, same material,
// not within five links -- are counted per higher voxel, the counts scanned, and each
// higher voxel writes its pairs in ascending order of the lower one, so the list in
// attPair comes out sorted with no sort pass.  The five-link test is exact here
// because links are only made later in this pass (a link only shortens paths).  One
// lane then takes the pairs in order, as the reference's threads would if they ran one
// after another: a pair within five links (counting the links made before it) is
// skipped, the dominant axis of the offset in the first voxel's frame picks the faces,
// and the link is made when both faces are free.  A pair table too small for this step
// stops the simulation in place (SC_COL_GROW) until the host enlarges it.

// is_neighbor (VX3_VoxelyzeKernel.cu:651-680): t is reachable from a within depth links,
// never stepping straight back over the link just taken
__device__ bool linkedWithin(const DevBatch<real>& b, size_t V, int a, int t, int depth) {
    if (a == t) return true;
    int node[6], skip[6], next[6];
    int top = 0;
    node[0] = a; skip[0] = -1; next[0] = 0;
    while (top >= 0) {
        if (next[top] >= 6 || top >= depth) { top--; continue; }
        const int u = node[top], d = next[top]++;
        if (d == skip[top] || !(b.vflags[u] & (1 << (VF_LINKMASK_SHIFT + d)))) continue;
        const int w = b.nbr[d * V + u];
        if (w == t) return true;
        top++;
        node[top] = w; skip[top] = d ^ 1; next[top] = 0;
    }
    return false;
}

// the candidate test of one (higher, lower) pair of attachable voxels
__device__ __forceinline__ bool attachCandidate(const DevBatch<real>& b, size_t V, double nom, int v1, int v2) {
    if (b.vmat[v2] != b.vmat[v1]) return false;
    const double b1 = nom * (double)b.tfac[v1], b2 = nom * (double)b.tfac[v2];
    const double wd = ((b1 + b1 + b1) / 3.0f + (b2 + b2 + b2) / 3.0f) * 0.625;   // COLLISION_ENVELOPE_RADIUS
    const double dx = (double)b.px[v1] - (double)b.px[v2], dy = (double)b.py[v1] - (double)b.py[v2],
                 dz = (double)b.pz[v1] - (double)b.pz[v2];
    if (dx > wd || dx < -wd || dy > wd || dy < -wd || dz > wd || dz < -wd) return false;
    if (sqrt(dx * dx + dy * dy + dz * dz) > wd) return false;
    return !linkedWithin(b, V, v1, v2, 5);   // covers is_neighbor depth 1 (directly linked)
}

// block-wide exclusive scan of x (256 threads); returns this thread's offset, total in *tot
__device__ __forceinline__ int blockScan256(int x, int* part, int* tot) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int incl = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) part[w] = incl;
    __syncthreads();
    int before = incl - x;
    for (int i = 0; i < w; i++) before += part[i];
    const int t = part[0] + part[1] + part[2] + part[3];
    __syncthreads();
    *tot = t;
    return before;
}

__global__ void __launch_bounds__(256) k_attach(DevBatch<real> b) {
    const int s = blockIdx.x;
    if (s >= b.nSims || !simRunning(b, s)) return;
    SimCtl& sc = b.sims[s];
    if (!(sc.flags & SC_ATTACH) || (sc.flags & (SC_DIVERGED_PENDING | SC_COL_OVERFLOW | SC_LINKS_DONE))) return;
    __shared__ int part[4];
    __shared__ int nc, np;
    const int tid = threadIdx.x;
    const size_t V = (size_t)b.Vcap;
    int* cand = b.attCand + sc.at0;   // attachable voxels (sim-local), then per candidate its pair offset
    int* poff = b.attCand + sc.at0 + sc.nv;
    if (tid == 0) { nc = 0; np = 0; }
    __syncthreads();
    for (int base = 0; base < sc.nv; base += 256) {
        const int i = base + tid;
        bool ok = false;
        if (i < sc.nv) {
            const int v = sc.v0 + i, fl = b.vflags[v];
            const VoxMatDev& m = b.vmatTab[b.vmat[v]];
            ok = !(fl & (VF_INTERIOR | VF_REMOVED)) && (fl & VF_ATTACH) && m.sticky && !m.fixed;
        }
        int tot;
        const int off = nc + blockScan256(ok ? 1 : 0, part, &tot);
        if (ok) cand[off] = i;
        if (tid == 0) nc += tot;
        __syncthreads();
    }
    const int n = nc;
    const double nom = sc.voxSize;
    // pairs per higher candidate x, then their offsets (chunks of 256 candidates)
    for (int base = 0; base < n; base += 256) {
        const int x = base + tid;
        int c = 0;
        if (x < n) {
            const int v1 = sc.v0 + cand[x];
            for (int y = 0; y < x; y++) c += attachCandidate(b, V, nom, v1, sc.v0 + cand[y]);
        }
        int tot;
        const int off = np + blockScan256(c, part, &tot);
        if (x < n) poff[x] = off;
        if (tid == 0) np += tot;
        __syncthreads();
    }
    const int p = np;
    if (p > sc.apCap) {   // the pair table is too small: stand still until the host enlarges it
        if (tid == 0) { sc.apNeed = p; atomicOr(&sc.flags, SC_COL_GROW); }
        return;
    }
    unsigned long long* sorted = b.attPair + sc.ap0;
    for (int x = tid; x < n; x += 256) {
        const int v1 = sc.v0 + cand[x];
        int k = poff[x];
        for (int y = 0; y < x; y++)
            if (attachCandidate(b, V, nom, v1, sc.v0 + cand[y])) sorted[k++] = ((unsigned long long)cand[x] << 32) | (unsigned)cand[y];
    }
    __syncthreads();
    if (tid != 0) return;
    int nl = sc.nl;
    bool made = false;
    for (int k = 0; k < p; k++) {
        const int v1 = sc.v0 + (int)(sorted[k] >> 32), v2 = sc.v0 + (int)(sorted[k] & 0xffffffffu);
        if (linkedWithin(b, V, v1, v2, 5)) continue;
        // the offset to the second voxel in the first voxel's frame
        const Q4<real> o1(b.ow[v1], b.ox[v1], b.oy[v1], b.oz[v1]);
        const V3<real> e(b.px[v2] - b.px[v1], b.py[v2] - b.py[v1], b.pz[v2] - b.pz[v1]);
        const V3<real> ea = o1.rotInv(e);
        const real fx = fabs(ea.x), fy = fabs(ea.y), fz = fabs(ea.z);
        int axis;
        real c;
        if (fx >= fy && fx >= fz) { axis = 0; c = ea.x; }
        else if (fy >= fx && fy >= fz) { axis = 1; c = ea.y; }
        else { axis = 2; c = ea.z; }
        const bool rev = c < 0;   // the second voxel lies on the first one's negative side
        const int d1 = 2 * axis + (rev ? 1 : 0), d2 = d1 ^ 1;
        if ((b.vflags[v1] & (1 << (VF_LINKMASK_SHIFT + d1))) || (b.vflags[v2] & (1 << (VF_LINKMASK_SHIFT + d2)))) continue;
        if (nl >= sc.nlMax) { sc.flags |= SC_COL_OVERFLOW; break; }
        const int l = sc.l0 + nl++;
        const int vneg = rev ? v2 : v1, vpos = rev ? v1 : v2;   // VX3_Link device ctor (VX3_Link.cu:30-55)
        b.lneg[l] = vneg;
        b.lpos[l] = vpos;
        b.lmeta[l] = axis | (b.vmatTab[b.vmat[v1]].selfLink << 2);
        b.lflags[l] = LF_SMALL_ANGLE | LF_FRESH;   // VX3_Link::reset
        b.q2x[l] = 0; b.q2y[l] = 0; b.q2z[l] = 0;
        b.a1x[l] = 0; b.a1y[l] = 0; b.a1z[l] = 0;
        b.a2x[l] = 0; b.a2y[l] = 0; b.a2z[l] = 0;
        b.strain[l] = 0; b.maxStrain[l] = 0; b.strainOffset[l] = 0;
        b.lnew[l] = sc.safety;
        b.lsim[l] = s;
        const int dn = 2 * axis, dp = 2 * axis + 1;
        for (int q = 0; q < 6; q++) {   // no force from the new link until its first update
            b.F[(dn * 6 + q) * b.Fs + vneg] = 0;
            b.F[(dp * 6 + q) * b.Fs + vpos] = 0;
        }
        b.es[dn * V + vneg] = 0;
        b.es[dp * V + vpos] = 0;
        b.nbr[dn * V + vneg] = vpos;
        b.nbr[dp * V + vpos] = vneg;
        b.vflags[vneg] |= 1 << (VF_LINKMASK_SHIFT + dn);
        b.vflags[vpos] |= 1 << (VF_LINKMASK_SHIFT + dp);
        made = true;
    }
    if (made) {
        sc.nl = nl;
        sc.flags |= SC_TOPO_CHG;
    }
}

// Target contacts (handle_collision_attachment, VX3_VoxelyzeKernel.cu:682-722): each
// watched pair of a target and another voxel, not both material-fixed, whose centres
// are within 0.625 x the sum of their average base sizes counts one hit per step, seen
// from the non-target voxel, which with signals receives 100 at once.  Positions and
// thermal factors from the contact snapshot (this step's start).
__global__ void __launch_bounds__(256) k_hits(DevBatch<real> b) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= b.vHi) return;
    const int s = b.vsim[v];
    if (!simRunning(b, s)) return;
    const SimCtl& sc = b.sims[s];
    if (!(sc.flags & SC_TARGETS) || (sc.flags & (SC_DIVERGED_PENDING | SC_COL_OVERFLOW))) return;
    const int fl = b.vflags[v];
    if (fl & (VF_INTERIOR | VF_TARGET)) return;
    const VoxMatDev& mv = b.vmatTab[b.vmat[v]];
    const double nom = sc.voxSize;
    const double bv = nom * (double)b.ctf[v];
    int n = 0;
    for (int k = 0; k < b.nbCount[v]; k++) {
        const int u = b.nbList[b.nbStart[v] + k];
        if (!(b.vflags[u] & VF_TARGET)) continue;
        if (mv.fixed && b.vmatTab[b.vmat[u]].fixed) continue;
        const double bu = nom * (double)b.ctf[u];
        const double wd = ((bv + bv + bv) / 3.0f + (bu + bu + bu) / 3.0f) * 0.625;   // baseSizeAverage, COLLISION_ENVELOPE_RADIUS
        const double dx = (double)b.cpx[v] - (double)b.cpx[u], dy = (double)b.cpy[v] - (double)b.cpy[u],
                     dz = (double)b.cpz[v] - (double)b.cpz[u];
        if (dx > wd || dx < -wd || dy > wd || dy < -wd || dz > wd || dz < -wd) continue;
        if (sqrt(dx * dx + dy * dy + dz * dz) > wd) continue;
        n++;
        if (sc.flags & SC_SIGNALS) receiveSignal(b, (size_t)b.Vcap, v, 100, sc.t, true);
    }
    if (n) atomicAdd(&b.sims[s].hits, n);
}

// Signals, second half: every voxel takes the signals its linked neighbours posted this
// step (the receiveSignal calls of their propagateSignal), in link-direction order.
__global__ void __launch_bounds__(256) k_signals(DevBatch<real> b) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= b.vHi) return;
    const int s = b.vsim[v];
    if (!simRunning(b, s)) return;
    const SimCtl& sc = b.sims[s];
    if (!(sc.flags & SC_SIGNALS) || (sc.flags & (SC_DIVERGED_PENDING | SC_COL_OVERFLOW))) return;
    const int fl = b.vflags[v];
    if (fl & VF_REMOVED) return;
    const size_t V = (size_t)b.Vcap;
    for (int d = 0; d < 6; d++) {
        if (!(fl & (1 << (VF_LINKMASK_SHIFT + d)))) continue;
        const int nb = b.nbr[d * V + v];
        const double val = b.sg[6 * V + nb];
        if (val > 0) receiveSignal(b, V, v, val, b.sg[7 * V + nb], false);
    }
}

// Block-wide deterministic FP64 center of mass of measured voxels (VX3_VoxelyzeKernel.cu:477-493).
__device__ void blockCoM(const DevBatch<real>& b, const SimCtl& sc, double out[3]) {
    __shared__ double red[4][256];
    double sx = 0, sy = 0, sz = 0, sm = 0;
    for (int v = sc.v0 + threadIdx.x; v < sc.v0 + sc.nv; v += blockDim.x) {
        const VoxMatDev& m = b.vmatTab[b.vmat[v]];
        if (!m.isMeasured) continue;
        const double mass = m.mass;
        sx += (double)b.px[v] * mass; sy += (double)b.py[v] * mass; sz += (double)b.pz[v] * mass;
        sm += mass;
    }
    red[0][threadIdx.x] = sx; red[1][threadIdx.x] = sy; red[2][threadIdx.x] = sz; red[3][threadIdx.x] = sm;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    const double tot = red[3][0];
    for (int k = 0; k < 3; k++) out[k] = tot == 0 ? 0.0 : (1.0 / tot) * red[k][0];
    __syncthreads();
}

// VoxCAD's per-step statistics for its stop conditions and equilibrium mode
// (CVX_Sim::TimeStep, UpdateStats, KineticEDecreasing; VoxCAD/src/VX1/VX_Sim.cpp:862-891,
// 1055-1160, 1833-1837): with equilibrium mode, all motion is zeroed when the kinetic
// energy of the last three steps is below that of the three before; then the total
// kinetic energy, the total strain energy and the largest voxel speed times dt are
// pushed into 500-step histories, newest first, -1 where the statistic is not needed.
// Sums run in double over float terms, as in the reference; the order of the sums is
// the block's, so they can differ from a sequential sum in the last bits.
__device__ void voxcadStats(const DevBatch<real>& b, SimCtl& sc, int s) {
    __shared__ double red[3][256];
    __shared__ int zero;
    double* H = b.hist + (size_t)s * 3 * VX3_HIST;   // [0, 500) kinetic, then total, then max move
    const bool kin = sc.scType == 4 || sc.scType == 5 || sc.eqMode;
    const bool strainE = sc.scType == 4;
    const bool vel = sc.scType == 6 || (sc.flags & SC_COLLISIONS);
    if (threadIdx.x == 0) {
        int z = 0;
        if (sc.eqMode) {
            double k[6];
            for (int i = 0; i < 6; i++) k[i] = H[(sc.histHead + i) % VX3_HIST];
            z = k[0] + k[1] + k[2] < k[3] + k[4] + k[5] &&
                !(k[0] == 0 || k[1] == 0 || k[2] == 0 || k[3] == 0 || k[4] == 0 || k[5] == 0);
        }
        zero = z;
    }
    __syncthreads();
    double ke = 0, v2 = 0, se = 0;
    for (int v = sc.v0 + threadIdx.x; v < sc.v0 + sc.nv; v += blockDim.x) {
        if (zero) {   // CVX_Voxel::haltMotion
            b.mx[v] = 0; b.my[v] = 0; b.mz[v] = 0;
            b.ax[v] = 0; b.ay[v] = 0; b.az[v] = 0;
        }
        const VoxMatDev& m = b.vmatTab[b.vmat[v]];
        const double lx = b.mx[v], ly = b.my[v], lz = b.mz[v];
        const double ax = b.ax[v], ay = b.ay[v], az = b.az[v];
        // CVX_Voxel::kineticEnergy (old/VX_Voxel.h:104)
        ke += (float)(0.5 * (m.massInverse * (lx * lx + ly * ly + lz * lz) + m.momentInertiaInverse * (ax * ax + ay * ay + az * az)));
        // velocity().Length2() (old/VX_Voxel.h:100)
        const double vx = lx * m.massInverse, vy = ly * m.massInverse, vz = lz * m.massInverse;
        v2 = fmax(v2, vx * vx + vy * vy + vz * vz);
    }
    if (strainE) {
        for (int l = sc.l0 + threadIdx.x; l < sc.l0 + sc.nl; l += blockDim.x) {
            // CVX_Link::strainEnergy (old/VX_Link.cpp:243-249) from the loads the link pass
            // left in the slots of its two voxels
            const int meta = b.lmeta[l], axis = meta & 3;
            const LinkMatDev& mat = b.lmatTab[meta >> 2];
            const int vn = b.lneg[l], vp = b.lpos[l], dn = 2 * axis, dp = 2 * axis + 1;
            const size_t FS = b.Fs;
            const double fNx = b.F[(dn * 6 + 0) * FS + vn];
            const double mNx = b.F[(dn * 6 + 3) * FS + vn], mNy = b.F[(dn * 6 + 4) * FS + vn], mNz = b.F[(dn * 6 + 5) * FS + vn];
            const double mPy = b.F[(dp * 6 + 4) * FS + vp], mPz = b.F[(dp * 6 + 5) * FS + vp];
            se += (float)(fNx * fNx / (2.0f * mat.a1) + mNx * mNx / (2.0 * mat.a2) +
                          (mNz * mNz - mNz * mPz + mPz * mPz) / (3.0 * mat.b3) + (mNy * mNy - mNy * mPy + mPy * mPy) / (3.0 * mat.b3));
        }
    }
    red[0][threadIdx.x] = ke; red[1][threadIdx.x] = v2; red[2][threadIdx.x] = se;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + w]);
            red[2][threadIdx.x] += red[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int head = (sc.histHead + VX3_HIST - 1) % VX3_HIST;
        const double dtD = sc.dtFixed > 0 ? sc.dtFixed : sc.dtFrac * sc.optDt;
        H[head] = kin ? red[0][0] : -1.0;
        H[VX3_HIST + head] = kin && strainE ? red[0][0] + red[2][0] : -1.0;
        H[2 * VX3_HIST + head] = vel ? sqrt(red[1][0]) * dtD : -1.0;
        sc.histHead = head;
    }
    __syncthreads();
}

// CVX_Sim::StopConditionMet (VoxCAD/src/VX1/VX_Sim.cpp:1015-1055) after `steps` steps at time sc.t
__device__ bool voxcadStopMet(const DevBatch<real>& b, const SimCtl& sc, int s, int steps) {
    if (steps < 2 * VX3_HIST) return false;
    const double* H = b.hist + (size_t)s * 3 * VX3_HIST;
    auto at = [&](int q, int age) { return H[q * VX3_HIST + (sc.histHead + age) % VX3_HIST]; };
    const double inv = 1.0 / (float)sc.nv;
    const double thr = 0.001 * sc.scValue;
    const int jump = VX3_HIST / 10;
    switch (sc.scType) {
    case 1: return steps > (int)(sc.scValue + 0.5);
    case 2: return sc.t > sc.scValue;
    case 3: return sc.t > sc.tempPeriod * sc.scValue;
    case 4: {
        const double ini = at(1, 0);
        for (int i = jump; i < VX3_HIST; i += jump) {
            if (at(1, i) == -1) return false;
            if (fabs(at(1, i) - ini) * inv > thr) return false;
        }
        return true;
    }
    case 5:
        for (int i = 0; i < VX3_HIST; i += jump) {
            if (at(0, i) == -1) return false;
            if (at(0, i) * inv > thr) return false;
        }
        return true;
    case 6:
        for (int i = 0; i < VX3_HIST; i += jump) {
            if (at(2, i) == -1) return false;
            if (at(2, i) > thr) return false;
        }
        return true;
    default: return false;
    }
}

// The secondary experiment at the end of a step (VX3_VoxelyzeKernel.cu:336-350):
// removeVoxels (:365-397) takes out the voxels of each material whose removal time is
// positive and earlier than the step's start time, with their links (gone from both
// voxels' masks; surfaces follow at the next step's start); then, once the start time
// passes the reset time, saveInitialPosition (:155-160) stores the current positions.
// Returns true when the initial center of mass is to be reset too.
__device__ bool secondaryExperiment(const DevBatch<real>& b, SimCtl& sc) {
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    for (int v = sc.v0 + threadIdx.x; v < sc.v0 + sc.nv; v += blockDim.x) {
        const double t = b.rt[v];   // the voxel material's removal time, 0 = never
        if (t > 0 && t < sc.t && !(b.vflags[v] & VF_REMOVED)) {
            atomicOr(&b.vflags[v], VF_REMOVED);
            any = 1;
        }
    }
    __syncthreads();
    if (any) {
        if (threadIdx.x == 0) atomicOr(&sc.flags, SC_TOPO_CHG);
        for (int l = sc.l0 + threadIdx.x; l < sc.l0 + sc.nl; l += blockDim.x) {
            const int lf = b.lflags[l];
            if (lf & LF_GONE) continue;
            const int vn = b.lneg[l], vp = b.lpos[l];
            if (!((b.vflags[vn] | b.vflags[vp]) & VF_REMOVED)) continue;
            const int axis = b.lmeta[l] & 3;
            b.lflags[l] = lf | LF_GONE;
            atomicAnd(&b.vflags[vn], ~(1 << (VF_LINKMASK_SHIFT + 2 * axis)));
            atomicAnd(&b.vflags[vp], ~(1 << (VF_LINKMASK_SHIFT + 2 * axis + 1)));
        }
        __syncthreads();
        // the Poisson strain the next link pass uses follows the voxels' remaining links
        // (CVX_Voxel::strain is evaluated lazily there), so the value the voxel pass
        // stored for it is recomputed
        if (sc.flags & SC_POISSON)
            for (int v = sc.v0 + threadIdx.x; v < sc.v0 + sc.nv; v += blockDim.x)
                if (!(b.vflags[v] & VF_REMOVED)) poissonStrain(b, v);
        __syncthreads();
    }
    const bool reinit = !sc.reinitDone && sc.reinitAfter < sc.t;
    if (reinit) {
        const size_t V = (size_t)b.Vcap;
        for (int v = sc.v0 + threadIdx.x; v < sc.v0 + sc.nv; v += blockDim.x) {
            b.ip[0 * V + v] = b.px[v]; b.ip[1 * V + v] = b.py[v]; b.ip[2 * V + v] = b.pz[v];
        }
    }
    __syncthreads();
    return reinit;
}

// computeTargetCloseness (VX3_VoxelyzeKernel.cu:545-562) over the simulation's target
// voxels: pairs closer than closeR are counted, 1/distance summed over all pairs (the
// block's partial sums, so the last bits of the sum can differ from a sequential one).
__device__ void blockCloseness(const DevBatch<real>& b, const SimCtl& sc, double& closeness, int& pairs) {
    __shared__ double rs[256];
    __shared__ int rc[256];
    double ret = 0;
    int cnt = 0;
    const int* tg = b.tgtList + sc.tg0;
    for (int i = threadIdx.x; i < sc.nTg; i += blockDim.x) {
        const int vi = tg[i];
        const double xi = b.px[vi], yi = b.py[vi], zi = b.pz[vi];
        for (int j = i + 1; j < sc.nTg; j++) {
            const int vj = tg[j];
            const double dx = (double)b.px[vj] - xi, dy = (double)b.py[vj] - yi, dz = (double)b.pz[vj] - zi;
            const double distance = sqrt(dx * dx + dy * dy + dz * dz);
            if (distance < sc.closeR) cnt++;
            ret += 1 / distance;
        }
    }
    rs[threadIdx.x] = ret;
    rc[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { rs[threadIdx.x] += rs[threadIdx.x + w]; rc[threadIdx.x] += rc[threadIdx.x + w]; }
        __syncthreads();
    }
    closeness = rs[0];
    pairs = rc[0];
    __syncthreads();
}

// Per-simulation bookkeeping after the voxel pass (doTimeStep tail, VX3_VoxelyzeKernel.cu:241,279,314-334,352,
// and the stop test the driver loop makes before the next step, VX3_SimulationManager.cu:62-64).
__global__ void __launch_bounds__(256) k_control(DevBatch<real> b) {
    __shared__ double sQ[VX3_DEV_MT_MAX + 1];
    const int s = blockIdx.x;
    if (s == 0 && threadIdx.x == 0) *b.colAny = 0;   // every rebuild pass of this step is done
    if (s >= b.nSims) return;
    SimCtl& sc = b.sims[s];
    if (!simRunning(b, s)) return;
    const bool diverged = (sc.flags & SC_DIVERGED_PENDING) != 0;
    const int steps = sc.steps + 1;
    const bool sample = !diverged && sc.cycleStep > 0 && steps % sc.cycleStep == 0;
    double com[3];
    if (sample) blockCoM(b, sc, com);
    double closeness = 0;
    int closePairs = 0;
    if (sample && sc.closeR > 0) blockCloseness(b, sc, closeness, closePairs);
    if (!diverged && !(sc.flags & SC_COL_OVERFLOW) && (sc.scType >= 4 || sc.eqMode)) voxcadStats(b, sc, s);
    bool reinit = false;
    if ((sc.flags & SC_SECONDARY) && !diverged && !(sc.flags & SC_COL_OVERFLOW)) reinit = secondaryExperiment(b, sc);
    if (threadIdx.x != 0) return;
    sc.steps = steps;
    sc.flags &= ~(SC_REBUILD | SC_LINKS_DONE);
    if (sc.flags & SC_TOPOCOL) {   // a change made during this step: nearby lists regenerated at the next
        sc.flags &= ~SC_NEARDIRTY;
        if (sc.flags & SC_TOPO_CHG) {
            sc.flags = (sc.flags & ~SC_TOPO_CHG) | SC_NEARDIRTY;
            sc.nearUsed = 0;
        }
    }
    // the next step rebuilds its contact lists after new nearby lists, or at the fixed period
    if ((sc.flags & SC_COLLISIONS) && ((sc.flags & SC_NEARDIRTY) || (sc.colInterval > 0 && steps % sc.colInterval == 0)))
        sc.flags |= SC_REBUILD;
    if (sc.flags & SC_COL_OVERFLOW) {
        sc.state = 4;
        return;
    }
    if (diverged) {
        sc.state = 2;
        return;
    }
    if (sample) {
        sc.angleSamples++;
        for (int k = 0; k < 3; k++) { sc.comH0[k] = sc.comH1[k]; sc.comH1[k] = sc.com[k]; sc.com[k] = com[k]; }
        const double* A = sc.comH0; const double* B = sc.comH1; const double* C = sc.com;
        const bool bc = B[0] == C[0] && B[1] == C[1] && B[2] == C[2];
        const bool ab = A[0] == B[0] && A[1] == B[1] && A[2] == B[2];
        if (bc || ab || sc.angleSamples < 3) sc.recentAngle = 0;
        else {
            double d1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, d2[3] = {C[0] - B[0], C[1] - B[1], C[2] - B[2]};
            double dot = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
            double l1 = sqrt((A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]));
            double l2 = sqrt((B[0] - C[0]) * (B[0] - C[0]) + (B[1] - C[1]) * (B[1] - C[1]) + (B[2] - C[2]) * (B[2] - C[2]));
            sc.recentAngle = acos(dot / (l1 * l2));
        }
        if (sc.closeR > 0) { sc.closeness = closeness; sc.closePairs = closePairs; }
    }
    if (reinit) {   // InitializeCenterOfMass: the last sampled center of mass
        for (int k = 0; k < 3; k++) sc.com0[k] = sc.com[k];
        sc.reinitDone = 1;
    }
    sc.t += (double)sc.dt;
    sc.prevDt = sc.dt;
    if (sc.maxSteps > 0 && steps >= sc.maxSteps) { sc.state = 1; return; }
    if (sc.stopLen > 0 &&
        mathTreeEval(b.tokens + sc.stopOff, sc.stopLen, sc.com[0], sc.com[1], sc.com[2], (double)sc.hits, sc.t, sc.recentAngle,
                     sc.closeness, sc.closePairs,
                     sc.nv, sQ) > 0)
        sc.state = 1;
    else if (sc.scType && voxcadStopMet(b, sc, s, steps))
        sc.state = 1;
}

// Recommended time step of each newly admitted simulation: max over its links of
// k/min(m1, m2) (VX3_VoxelyzeKernel.cu:184-217), wave shuffles + LDS in FP64.
__device__ __forceinline__ double waveMax(double x) {
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    return x;
}
__global__ void __launch_bounds__(256) k_dt(DevBatch<real> b, int s0, int s1) {
    const int s = s0 + blockIdx.x;
    if (s >= s1) return;
    SimCtl& sc = b.sims[s];
    if (sc.state != 0 || sc.optDt != 0.0) return;
    __shared__ double part[8];
    double mx = 0.0;
    const float size = (float)sc.voxSize;
    const real nom = (real)sc.voxSize;
    for (int l = sc.l0 + threadIdx.x; l < sc.l0 + sc.nl; l += blockDim.x) {
        const int vn = b.lneg[l], vp = b.lpos[l];
        const LinkMatDev& mat = b.lmatTab[b.lmeta[l] >> 2];
        const VoxMatDev& mn = b.vmatTab[b.vmat[vn]];
        const VoxMatDev& mp = b.vmatTab[b.vmat[vp]];
        float k;
        if (mat.nu == 0.0f) k = mat.a1;
        else {
            const int axis = b.lmeta[l] & 3;
            float an, sn, ap, sp;
            const float psN[3] = {b.psx[vn], b.psy[vn], b.psz[vn]}, psP[3] = {b.psx[vp], b.psy[vp], b.psz[vp]};
            transverseOf(psN, axis, size, mn.nu == 0.0f, an, sn);
            transverseOf(psP, axis, size, mp.nu == 0.0f, ap, sp);
            const float area = 0.5f * (an + ap);
            const double rl = 0.5 * ((double)nom * (double)b.tfac[vn] + (double)nom * (double)b.tfac[vp]);
            k = (float)(mat.eHat * area / ((b.strain[l] + 1) * rl));
        }
        const double m1 = mn.mass, m2 = mp.mass;
        const double f2 = k / (m1 < m2 ? m1 : m2);
        if (f2 > mx) mx = f2;
    }
    if (sc.nl == 0) {
        for (int v = sc.v0 + threadIdx.x; v < sc.v0 + sc.nv; v += blockDim.x) {
            const VoxMatDev& m = b.vmatTab[b.vmat[v]];
            const double f2 = m.E * m.nomSize / m.mass;
            if (f2 > mx) mx = f2;
        }
    }
    mx = waveMax(mx);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) part[wid] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, part[w]);
        double opt = m <= 0.0f ? 0.0 : 1.0f / (6.283185f * sqrt(m));
        if (opt < 1e-10) opt = 1e-10;
        sc.optDt = opt;
        if (!(sc.dt > 0)) sc.dt = (float)(sc.dtFrac * opt);   // doTimeStep: dt = DtFrac * OptimalDt (float)
        if (!(sc.prevDt > 0)) sc.prevDt = sc.dt;
        sc.cycleStep = sc.tempPeriod / (double)sc.dt >= 1.0 ? (int)(sc.tempPeriod / (double)sc.dt) : 0;
    }
}

}  // namespace VX_NS
}  // namespace vx3
