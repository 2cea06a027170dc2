// vx3_resident.inl -- persistent multi-step kernel: one workgroup advances one
// simulation for many steps with its working set on chip.
//
// Each thread owns VPT voxels and, for each of the three link axes, LPA links, and
// keeps their mutable state in registers for the whole launch.  What other threads
// need crosses LDS:
//   pose    position + orientation + thermal factor + Poisson strain per voxel,
//           written by the voxel owner, read by the link owners;
//   loads   the two local end loads of the links of ONE axis at a time: the link
//           pass runs axis by axis (X, Y, Z) and after each axis the voxel owners add
//           that axis' two directions to their running sums.  Summing X_POS, X_NEG,
//           Y_POS, ... in this order is exactly CVX_Voxel::force()/moment()
//           (Voxelyze3/src/old/VX_Voxel.cpp:232-269), and the load buffer holds one
//           axis instead of all links;
//   tables  the simulation's voxel and link material constants.
// A step is: 3 x (link pass of one axis | barrier | partial gather | barrier),
// voxel integration + per-simulation control, barrier.
// Global memory is read once at launch start and written once at the end in the
// stream path's layout, so the two paths can hand simulations to each other.
//
// Collisions (COL): a simulation with self-collision keeps its contact partner lists
// in LDS (sNb, ascending per voxel, the stream layout's order) across the launch.  At a
// step that the staleness rule (a surface voxel moved more than recalcDist since the
// last rebuild; checked after each voxel update) or the fixed period marks, the lists
// are rebuilt before the link pass: surface voxels are counting-sorted into LDS hash
// buckets of their cell (edge = watch radius), and every surface voxel tests the
// voxels of its 27 neighbouring cells against the CPU engine's pair rule (watch
// radius, link-hop exclusion), as the per-step kernels do.  After the link passes the
// load buffer holds a snapshot of positions and momenta at the step's start, from which
// the voxel pass computes each contact (the thermal factor of the next step is written
// only after every voxel has used this step's).  Lists that do not fit the LDS table
// stop the simulation in place (SC_COL_GROW); the host then moves it to the per-step
// kernels.
//
// Tiled (TILED): a simulation too large for one workgroup runs on several, one per tile
// (TileDesc).  A tile holds in LDS the poses of the voxels it owns and of its halo (the
// other ends of its owned voxels' links), computes every link with an owned end and
// integrates only its owned voxels; a link between two tiles is computed by both from
// the same inputs, bit for bit, so no loads cross tiles.  After each step's voxel pass
// a tile writes its boundary voxels' poses to the simulation's exchange area (one of
// two, by step parity), arrives at the simulation's barrier and, once every tile has,
// reads its halo back.  The centre-of-mass partial sums and a divergence word travel
// through the same barrier, so every tile takes the same control decisions.  A link
// that diverges in one tile is seen by the others only after the step's voxel pass:
// the tiles then leave without writing anything back (global state is untouched until
// the end of a launch), and a repair launch right after runs those simulations again
// from the same start to the link pass of that step, where the one-workgroup kernel
// stops.  A tile that waits at a barrier longer than TileArgs::spin wall-clock ticks
// fails the simulation instead of waiting on.  A simulation of one tile (a one-workgroup
// simulation placed in a tiled launch) has no barrier or exchange and steps exactly as
// the one-workgroup kernel does.
//
// Included by the per-precision kernel units after vx3_kernels.inl (VX_NS, real).

namespace vx3 {
namespace VX_NS {

#define VX3_RES_PHASES 64   // distinct actuation phases evaluated once per step instead of per voxel
#ifndef VX3_RES_VMATS
#define VX3_RES_VMATS 8      // voxel materials a resident sim may use
#define VX3_RES_LMATS 16     // link material variants a resident sim may use
#endif

#define VX3_RES_PAIRS 8192   // contact partner entries (both directions) a resident sim may hold

template <int BS, int VPT, int LPA, bool COL = false, bool TILED = false>
struct ResidentShape {
    static constexpr int VMAX = BS * VPT;   // voxels per sim
    static constexpr int LAX = BS * LPA;    // links per axis per sim
    static constexpr size_t tables = VX3_RES_VMATS * sizeof(VoxMatDev) + VX3_RES_LMATS * sizeof(LinkMatDev);
    static constexpr size_t bytes = tables
                                  + (size_t)VMAX * 7 * sizeof(real)      // pose
                                  + (size_t)LAX * 12 * sizeof(real)      // loads of one axis
                                  + (size_t)VMAX * 4 * sizeof(float)     // thermal factor, Poisson strain
                                  + (size_t)LAX * 6 * sizeof(float)      // end strains (axis, end, slot)
                                  + (COL ? (size_t)VX3_RES_PAIRS * 2 : 0)   // contact partners
                                  + (TILED ? (size_t)VMAX * 4 : 0);         // exchange voxel per slot
    // setup tables alias the load buffer: voxel -> slot per direction, and (axis, slot) -> link
    static_assert((size_t)VMAX * 6 * 2 + (size_t)LAX * 3 * 2 <= (size_t)LAX * 12 * sizeof(real), "setup alias");
    // with collisions the load buffer also holds the rebuild's buckets (2 ints per bucket, one
    // bucket per voxel) and sorted voxels, and later the step-start snapshot (6 reals per voxel)
    static_assert(!COL || (size_t)VMAX * 10 <= (size_t)LAX * 12 * sizeof(real), "rebuild alias");
    static_assert(!COL || (size_t)VMAX * 6 <= (size_t)LAX * 12, "snapshot alias");
    static_assert((VMAX & (VMAX - 1)) == 0, "bucket count is a power of two");
};

// block-wide exclusive scan of one int per thread; *total receives the sum.  part holds
// BS/64 ints.  Ends with the block synchronized.
template <int BS>
__device__ __forceinline__ int resBlockScan(int x, int* part, int* total) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int incl = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) part[w] = incl;
    __syncthreads();
    int before = incl - x, t = 0;
    for (int i = 0; i < BS / 64; i++) {
        if (i < w) before += part[i];
        t += part[i];
    }
    __syncthreads();
    *total = t;
    return before;
}

// contact force of a resident simulation's voxel from its partners (sim-local indices in
// nb[start, start + count), ascending), from the step-start snapshot snap = [6][VMAX]
// (position, momentum) and this step's thermal factors tf
struct ResContacts {
    const DevBatch<real>* b;
    const real* snap;
    const float* tf;
    const unsigned short* nb;
    int vmax, v, v0, start, count;
    real nom;
    template <typename T> __device__ __forceinline__ void subtract(V3<T>& force) const {
        for (int k = 0; k < count; k++) {
            const int u = nb[start + k];
            const int i = v < u ? v : u, j = v < u ? u : v;
            float fx, fy, fz;
            pairContact(b->vmatTab[b->vmat[v0 + i]], b->vmatTab[b->vmat[v0 + j]], snap[i], snap[vmax + i], snap[2 * vmax + i],
                        snap[j], snap[vmax + j], snap[2 * vmax + j], snap[3 * vmax + i], snap[4 * vmax + i], snap[5 * vmax + i],
                        snap[3 * vmax + j], snap[4 * vmax + j], snap[5 * vmax + j], tf[i], tf[j], nom, fx, fy, fz);
            if (v != i) { fx = -fx; fy = -fy; fz = -fz; }
            force.x -= (T)fx; force.y -= (T)fy; force.z -= (T)fz;
        }
    }
};

// the CPU pair rule between surface voxels v and u (sim-local) of a resident simulation,
// positions from the LDS pose cache (colWatched of the per-step kernels)
__device__ __forceinline__ bool resWatched(const DevBatch<real>& b, const SimCtl& sc, const real* sPx, const real* sPy,
                                           const real* sPz, int v, int u) {
    const real dx = sPx[v] - sPx[u], dy = sPy[v] - sPy[u], dz = sPz[v] - sPz[u];
    if (dx * dx + dy * dy + dz * dz > (real)sc.colThreshSq) return false;
    const int gv = sc.v0 + v, gu = sc.v0 + u;
    const int li = abs(b.vijk[3 * gv] - b.vijk[3 * gu]) + abs(b.vijk[3 * gv + 1] - b.vijk[3 * gu + 1]) +
                   abs(b.vijk[3 * gv + 2] - b.vijk[3 * gu + 2]);
    if (li > sc.colNearHops) return true;
    return !nearbyOf(b, sc, gv, gu);
}

__device__ __forceinline__ void ldsCopyWords(int* dst, const int* src, int words, int tid, int bs) {
    for (int i = tid; i < words; i += bs) dst[i] = src[i];
}

// Data that crosses tiles is written and read with relaxed agent-scope atomics, which the
// compiler issues coherent across the device's L2 caches; so no fence has to write back
// or invalidate a whole L2 per step.  Ordering comes from waiting for every outstanding
// memory operation (s_waitcnt 0: the stores have reached their point of coherence)
// before the arrival flag is written.
template <typename T> __device__ __forceinline__ void xput(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ T xget(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The simulation's barrier, in two halves.  Each tile has an arrival flag: one thread
// writes it, after every thread of the tile has written its exchange data and waited for
// it (s_waitcnt 0) and the workgroup has synchronized; later the first wave waits, one
// flag per lane, until every tile's flag shows barrier number `gen` (0-based).  Flags
// rather than one counter: no read-modify-write queue on one address, and the wave's
// loads of all flags take one memory latency.  A repair launch counts from 1 << 24, so
// the flags its launch left need no clearing.  False when the wait exceeded `spin` ticks.
__device__ __forceinline__ int tileFlag(unsigned gen, int repair) { return (int)(gen + 1) + (repair ? (1 << 24) : 0); }
__device__ __forceinline__ void tileArrive(int* flag, unsigned gen, int repair) {
    __builtin_amdgcn_s_waitcnt(0);
    xput(flag, tileFlag(gen, repair));
}
// called by the 64 lanes of one wave; the same result in each
__device__ __forceinline__ bool tileWait(const int* flags, int n, unsigned gen, int repair, long long spin) {
    const int target = tileFlag(gen, repair);
    const long long t0 = wall_clock64();
    for (;;) {
        bool ready = true;
        for (int q = (int)(threadIdx.x & 63); q < n; q += 64) ready = ready && xget(flags + q) >= target;
        if (__all(ready)) return true;
        if (wall_clock64() - t0 > spin) return false;
        __builtin_amdgcn_s_sleep(1);
    }
}

template <int BS, int VPT, int LPA, bool COL, bool TILED = false>
__global__ void __launch_bounds__(BS) k_resident(DevBatch<real> b, const int* simList, int nSteps, TileArgs ta) {
    static_assert(!(COL && TILED), "tiled simulations have no collisions");
    typedef ResidentShape<BS, VPT, LPA, COL, TILED> Shape;
    constexpr int VMAX = Shape::VMAX, LAX = Shape::LAX;
    extern __shared__ __align__(16) unsigned char smem[];
    VoxMatDev* sVM = (VoxMatDev*)smem;
    LinkMatDev* sLM = (LinkMatDev*)(smem + VX3_RES_VMATS * sizeof(VoxMatDev));
    real* sPx = (real*)(smem + Shape::tables);
    real* sPy = sPx + VMAX; real* sPz = sPy + VMAX;
    real* sOw = sPz + VMAX; real* sOx = sOw + VMAX; real* sOy = sOx + VMAX; real* sOz = sOy + VMAX;
    real* sF = sOz + VMAX;                      // [12][LAX]
    float* sTf = (float*)(sF + 12 * LAX);
    float* sPsx = sTf + VMAX; float* sPsy = sPsx + VMAX; float* sPsz = sPsy + VMAX;
    float* sEs = sPsz + VMAX;                   // [(axis*2 + end)][LAX]
    unsigned short* sNb = (unsigned short*)(sEs + 6 * LAX);   // contact partners (COL)
    int* sXv = (int*)(sEs + 6 * LAX);   // TILED: per slot its exchange record (owned: another tile reads
                                        // it; halo: its owner's), else -1
    unsigned short* sVL = (unsigned short*)sF;  // [6][VMAX] slot of the link in direction d (setup only)
    unsigned short* sSL = sVL + 6 * VMAX;       // [3][LAX]  link of (axis, slot) (setup only)
    __shared__ int sRun, sDiv, sCnt[3], sExp[3], sExpAt[3];
    __shared__ double sQ[VX3_DEV_MT_MAX + 1];
    __shared__ double sRed[4][BS / 64];
    __shared__ double sCom[3], sComH0[3], sComH1[3], sAngle;   // center-of-mass history, owned by thread 0
    __shared__ double sPhase[VX3_RES_PHASES];   // distinct actuation phase offsets of the sim
    __shared__ float sPhT[VX3_RES_PHASES];      // their temperatures for the step being run
    __shared__ int sNph;                        // table size; -1: more than VX3_RES_PHASES (admission prevents it)
    __shared__ float sDm[VX3_RES_VMATS];        // CVX_Voxel::dampingMultiplier per voxel material (dt is fixed)
    __shared__ int sStale[2], sColOvf;          // COL: a rebuild is due (per step parity); the pair table overflowed
    __shared__ int sScan[BS / 64];
    __shared__ int sTile;                       // TILED: 1 = some tile diverged this step, 2 = a barrier timed out

    TileDesc td{};
    if (TILED) td = ta.tiles[blockIdx.x];
    const int s = TILED ? td.sim : simList[blockIdx.x];
    SimCtl& scg = b.sims[s];
    const int tid = threadIdx.x;
    // TILED: a simulation of one tile (a one-workgroup simulation run inside a tiled launch)
    // has no barrier and no exchange: it steps as the one-workgroup kernel does
    const bool multi = TILED && td.nTiles > 1;
    int divAt = -1;      // TILED: the launch's step at which a link diverged (a repair launch stops there)
    unsigned gen = 0;    // TILED: barriers passed in this launch
    if (scg.state != 0 || (scg.flags & SC_COL_GROW)) return;   // finished, or waiting to move to the per-step kernels
    if (TILED && ta.repair) {   // only simulations whose previous launch saw a divergence run again
        const int dw = ta.sync[td.cOff];
        if (!dw) return;
        divAt = dw - 1;
    }
    const int v0 = scg.v0, nv = scg.nv, l0 = scg.l0, nl = scg.nl;
    // the voxels this workgroup integrates, the LDS pose slots it fills, the links it computes
    const int nvOwn = TILED ? td.nOwn : nv, nvAll = TILED ? td.nOwn + td.nHalo : nv, nlT = TILED ? td.nLinks : nl;
    auto gvox = [&](int slot) { return v0 + (TILED ? ta.vox[td.vOff + slot] & 0x7fffffff : slot); };
    auto glink = [&](int j) { return l0 + (TILED ? ta.link[td.lOff + j] & 0x7fffffff : j); };
    const double dt = scg.dt;
    const float dtf = scg.dt;
    const real nom = (real)scg.voxSize;
    const float size = (float)scg.voxSize;
    const int flags0 = scg.flags;
    const bool poisson = (flags0 & SC_POISSON) != 0;
    const bool vary = (flags0 & SC_VARY) != 0;
    const bool expansion = (flags0 & SC_EXPANSION) != 0;
    const double tempAmp = scg.tempAmp, tempPeriod = scg.tempPeriod;
    const int cycleStep = scg.cycleStep, maxSteps = scg.maxSteps;
    const int stopLen = scg.stopLen, stopOff = scg.stopOff;
    const int mat0 = scg.mat0, nvm = scg.nvm, nlm = scg.nlm;
    const size_t VC = (size_t)b.Vcap;

    // ---- setup: material tables, link slots, voxel -> slot table
    for (int i = 0; i < nvm; i++)
        ldsCopyWords((int*)&sVM[i], (const int*)&b.vmatTab[b.simMats[mat0 + i]], sizeof(VoxMatDev) / 4, tid, BS);
    for (int i = 0; i < nlm; i++)
        ldsCopyWords((int*)&sLM[i], (const int*)&b.lmatTab[b.simMats[mat0 + nvm + i]], sizeof(LinkMatDev) / 4, tid, BS);
    for (int i = tid; i < 6 * VMAX; i += BS) sVL[i] = 0xFFFF;
    if (tid < 3) { sCnt[tid] = 0; sExp[tid] = 0; sExpAt[tid] = 0; }
    __syncthreads();
    // Links that took the expensive branches last step (large-angle frame, or a
    // relative rotation past toRotVec's series range, |a2v| >= ~0.098 rad) get the
    // lowest slots of their axis, so they share as few wavefronts as possible: a
    // wavefront runs a branch if any of its lanes takes it.  Which thread computes a
    // link does not change its result.
    auto expensive = [&](int g) {
        const real x = b.a2x[g], y = b.a2y[g], z = b.a2z[g];
        return !(b.lflags[g] & LF_SMALL_ANGLE) || x * x + y * y + z * z > (real)0.009;
    };
    for (int l = tid; l < nlT; l += BS) {
        const int g = glink(l);
        if (expensive(g)) atomicAdd(&sExp[b.lmeta[g] & 3], 1);
    }
    __syncthreads();
    // Within each class (axis, expensive or not) links take slots in link-list order:
    // expensive links [0, sExp), the others [sExp, count) (admission guarantees count
    // <= LAX).  Link order follows voxel order, so neighbouring lanes gather neighbouring
    // voxels' poses and load slots, which keeps LDS bank conflicts low.  Ranks come from
    // wave ballots per class and per-wave counts in LDS; sCnt[a] / sExpAt[a] run as the
    // class bases over the chunks of BS links.
    {
        constexpr int NW = BS / 64;
        __shared__ int sWc[NW][6];
        const int lane = tid & 63, w = tid >> 6;
        const unsigned long long below = (1ull << lane) - 1;
        for (int base = 0; base < nlT; base += BS) {
            const int l = base + tid;
            int ax = 0, cls = -1;
            bool ex = false;
            if (l < nlT) {
                const int g = glink(l);
                ax = b.lmeta[g] & 3;
                ex = expensive(g);
                cls = 2 * ax + (ex ? 0 : 1);
            }
            int rank = 0;
#pragma unroll
            for (int c = 0; c < 6; c++) {
                const unsigned long long m = __ballot(cls == c);
                if (lane == 0) sWc[w][c] = __popcll(m);
                if (cls == c) rank = __popcll(m & below);
            }
            __syncthreads();
            if (cls >= 0) {
                for (int i = 0; i < w; i++) rank += sWc[i][cls];
                const int slot = ex ? sExpAt[ax] + rank : sExp[ax] + sCnt[ax] + rank;
                const int g = glink(l);
                const int e = TILED ? ta.ends[td.lOff + l] : (b.lneg[g] - v0) | ((b.lpos[g] - v0) << 16);
                sSL[ax * LAX + slot] = (unsigned short)l;
                sVL[(2 * ax) * VMAX + (e & 0xFFFF)] = (unsigned short)slot;
                sVL[(2 * ax + 1) * VMAX + (e >> 16)] = (unsigned short)slot;
            }
            __syncthreads();
            if (tid < 3) {
                for (int i = 0; i < NW; i++) { sExpAt[tid] += sWc[i][2 * tid]; sCnt[tid] += sWc[i][2 * tid + 1]; }
            }
            __syncthreads();
        }
    }
    if (tid < 3) sCnt[tid] += sExp[tid];   // links of axis tid
    // 2*sqrt(m)*zeta/previousDt (VX_Voxel.h:131): the first step's previousDt is the dt of the
    // simulation's last step (which a caller-given dt may have set), every later one the launch's dt
    const float prevDt0 = scg.prevDt > 0 ? scg.prevDt : dtf;
    if (tid < nvm) sDm[tid] = 2 * sVM[tid].sqrtMass * sVM[tid].zetaInternal / (sVM[tid].fixed ? 0.0f : prevDt0);
    __syncthreads();
    auto localVmat = [&](int gidx) {
        int r = 0;
        for (int i = 0; i < nvm; i++) if (b.simMats[mat0 + i] == gidx) r = i;
        return r;
    };
    auto localLmat = [&](int gidx) {
        int r = 0;
        for (int i = 0; i < nlm; i++) if (b.simMats[mat0 + nvm + i] == gidx) r = i;
        return r;
    };
    // links: (axis a, k) -> slot tid + k*BS of axis a
    int lEnds[3][LPA];    // vneg | vpos << 16 (sim-local), -1 = no link
    int lMats[3][LPA];    // link mat | neg voxel mat << 8 | pos voxel mat << 16 (local)
    int lIdx[3][LPA];     // sim-local link index
    LinkState<real> ls[3][LPA];
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int k = 0; k < LPA; k++) {
            const int slot = tid + k * BS;
            lEnds[a][k] = -1;
            if (slot < sCnt[a]) {
                const int l = sSL[a * LAX + slot];
                const int g = glink(l);
                lIdx[a][k] = l;
                lEnds[a][k] = TILED ? ta.ends[td.lOff + l] : (b.lneg[g] - v0) | ((b.lpos[g] - v0) << 16);
                lMats[a][k] = localLmat(b.lmeta[g] >> 2) | (localVmat(b.vmat[b.lneg[g]]) << 8) | (localVmat(b.vmat[b.lpos[g]]) << 16);
                ls[a][k].pos2 = V3<real>(b.q2x[g], b.q2y[g], b.q2z[g]);
                ls[a][k].a1v = V3<real>(b.a1x[g], b.a1y[g], b.a1z[g]);
                ls[a][k].a2v = V3<real>(b.a2x[g], b.a2y[g], b.a2z[g]);
                ls[a][k].strain = b.strain[g];
                ls[a][k].maxStrain = b.maxStrain[g];
                ls[a][k].strainOffset = b.strainOffset[g];
                ls[a][k].flags = b.lflags[g];
            }
        }
    }
    // position and orientation live only in the LDS pose cache (their single copy);
    // momenta stay in registers
    // Actuation phases: the temperature depends on the voxel only through its phase
    // offset, and simulations use few distinct offsets.  Wave 0 collects them
    // (lane i keeps entry i), so each step evaluates sin once per distinct offset;
    // the values are those the per-voxel formula gives.
    if (tid < 64) {
        // compared as bit patterns: every value, NaN included, equals itself
        long long tab = 0;
        int n = 0;
        bool over = false;
        for (int base = 0; base < nvOwn && !over; base += 64) {
            const int v = base + tid;
            const long long ph = v < nvOwn ? __double_as_longlong(b.phase[gvox(v)]) : 0;
            bool pending = v < nvOwn;
            while (!over && __any(pending)) {
                const int src = __ffsll((unsigned long long)__ballot(pending)) - 1;   // lowest pending lane
                const long long cand = __shfl(ph, src, 64);
                if (!__any(tid < n && tab == cand)) {
                    if (n == VX3_RES_PHASES) over = true;
                    else {
                        if (tid == n) tab = cand;
                        n++;
                    }
                }
                if (ph == cand) pending = false;   // clears at least lane src
            }
        }
        if (tid < n) sPhase[tid] = __longlong_as_double(tab);
        if (tid == 0) sNph = over ? -1 : n;
    }
    __syncthreads();
    const int nph = sNph;
    if (nph < 0) {
        if (tid == 0) scg.state = 4;   // failed
        return;
    }
    V3<real> vm[VPT], va[VPT];
    int vf[VPT], vmat[VPT], vext[VPT];
    int vpi[VPT];         // index of the voxel's phase offset in sPhase
    int vl[VPT][3];       // slots of directions (2i, 2i+1) packed lo | hi << 16, 0xFFFF = none
    // CVX_Voxel::strain's per-axis case, fixed for the launch (topology and externals do
    // not change): bit i = axis i has links on both sides (their strains are averaged),
    // bit 3+i = axis i is in tension, i.e. has two links or one link and a fixed or
    // forced degree of freedom on that axis (VX_Voxel.cpp:299-333)
    int vpm[VPT];
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int v = tid + k * BS;
        vmat[k] = -1;
        if (v < nvOwn) {
            const int g = gvox(v);
            vm[k] = V3<real>(b.mx[g], b.my[g], b.mz[g]);
            va[k] = V3<real>(b.ax[g], b.ay[g], b.az[g]);
            vf[k] = b.vflags[g];
            vmat[k] = localVmat(b.vmat[g]);
            vext[k] = b.vext[g];
            const long long ph = __double_as_longlong(b.phase[g]);
            vpi[k] = 0;
            for (int i = 0; i < nph; i++)
                if (__double_as_longlong(sPhase[i]) == ph) vpi[k] = i;
#pragma unroll
            for (int i = 0; i < 3; i++) vl[k][i] = (int)sVL[(2 * i) * VMAX + v] | ((int)sVL[(2 * i + 1) * VMAX + v] << 16);
            const ExtDev* ext = vext[k] >= 0 ? &b.extTab[vext[k]] : nullptr;
            vpm[k] = 0;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const int n = ((vl[k][i] & 0xFFFF) != 0xFFFF) + (((vl[k][i] >> 16) & 0xFFFF) != 0xFFFF);
                const bool tension = (n == 2) || (ext && n == 1 && ((ext->dofFixed & (1 << i)) || ext->force[i] != 0));
                vpm[k] |= (n == 2 ? 1 << i : 0) | (tension ? 8 << i : 0);
            }
        }
    }
    __syncthreads();   // setup tables alias the load buffer: done with them
    if (TILED)
        for (int v = tid; v < VMAX; v += BS) sXv[v] = v < nvAll ? ta.xidx[td.vOff + v] : -1;
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int v = tid + k * BS;
        if (v < nvAll) {
            const int g = gvox(v);
            sPx[v] = b.px[g]; sPy[v] = b.py[g]; sPz[v] = b.pz[g];
            sOw[v] = b.ow[g]; sOx[v] = b.ox[g]; sOy[v] = b.oy[g]; sOz[v] = b.oz[g];
            sTf[v] = b.tfac[g];
            sPsx[v] = b.psx[g]; sPsy[v] = b.psy[g]; sPsz[v] = b.psz[g];
        }
    }
    if (poisson) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
#pragma unroll
            for (int k = 0; k < LPA; k++) {
                if (lEnds[a][k] < 0) continue;
                const int slot = tid + k * BS;
                const int vn = lEnds[a][k] & 0xFFFF, vq = lEnds[a][k] >> 16;
                const int gn = TILED ? b.lneg[glink(lIdx[a][k])] : v0 + vn, gq = TILED ? b.lpos[glink(lIdx[a][k])] : v0 + vq;
                sEs[(2 * a) * LAX + slot] = b.es[(2 * a) * VC + gn];
                sEs[(2 * a + 1) * LAX + slot] = b.es[(2 * a + 1) * VC + gq];
            }
        }
    }
    // ---- contact lists (COL): the stream layout's lists into LDS, last-rebuild positions into
    // registers, and the staleness rule for the launch's first step
    const bool colFirst = COL && (flags0 & SC_REBUILD) != 0;
    float lw[VPT][3];
    int nbS[VPT], nbC[VPT];
    int rebuilds = 0;
    bool colOvf = false;
    if (COL) {
        if (tid < 2) sStale[tid] = 0;
        if (tid == 0) sColOvf = 0;
        __syncthreads();
        const int nb0 = scg.nb0;
        const real recalcSq = (real)scg.colRecalcSq;
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            nbS[k] = nbC[k] = 0;
            lw[k][0] = lw[k][1] = lw[k][2] = 0;
            const int v = tid + k * BS;
            if (vmat[k] < 0 || (vf[k] & VF_INTERIOR)) continue;
            const int g = v0 + v;
            lw[k][0] = b.lastWatch[3 * g]; lw[k][1] = b.lastWatch[3 * g + 1]; lw[k][2] = b.lastWatch[3 * g + 2];
            if (colFirst) continue;
            const int st0 = b.nbStart[g] - nb0, c = b.nbCount[g];
            if (c > 0) {
                if (st0 < 0 || st0 + c > VX3_RES_PAIRS) { sColOvf = 1; continue; }
                for (int i = 0; i < c; i++) sNb[st0 + i] = (unsigned short)(b.nbList[nb0 + st0 + i] - v0);
                nbS[k] = st0;
                nbC[k] = c;
            }
            const real dx = sPx[v] - (real)lw[k][0], dy = sPy[v] - (real)lw[k][1], dz = sPz[v] - (real)lw[k][2];
            if (dx * dx + dy * dy + dz * dz > recalcSq) sStale[0] = 1;
        }
    }
    double t = scg.t;
    int steps = scg.steps;
    int angleSamples = scg.angleSamples;
    int state = 0;
    if (tid == 0) {
        sRun = 1; sDiv = 0;
        for (int q = 0; q < 3; q++) { sCom[q] = scg.com[q]; sComH0[q] = scg.comH0[q]; sComH1[q] = scg.comH1[q]; }
        sAngle = scg.recentAngle;
    }
    __syncthreads();

    if (COL && sColOvf) colOvf = true;   // lists larger than the LDS table: no step here
    // TILED: the second half of the simulation's barrier `gen`, whose arrival is pending:
    // wait for every tile, then (sampleNow) the tiles' centre-of-mass partial sums in tile
    // order into sRed[.][0], then (loadHalo) the halo poses into LDS.  1: some tile saw a
    // link diverge (leave; the repair launch takes over), 2: a tile did not arrive (the
    // simulation fails).
    bool pend = false;
    auto finishBarrier = [&](bool loadHalo, bool sampleNow) -> int {
        if (tid < 64) {   // the first wave; tile 0 of the simulation is workgroup blockIdx.x - td.tile
            int flag = 0;
            if (!tileWait(ta.sync + ta.flag0 + (blockIdx.x - td.tile), td.nTiles, gen, ta.repair, ta.spin)) flag = 2;
            else if (divAt < 0 && xget(ta.sync + td.cOff)) flag = 1;
            if (tid == 0) sTile = flag;
        }
        const int par = gen & 1;
        gen++;
        pend = false;
        __syncthreads();
        if (sTile == 2) {   // state stays as it was
            if (tid == 0) scg.state = 4;
            return 2;
        }
        if (sTile == 1) return 1;   // state untouched: the repair launch runs it to that step's link pass
        if (sampleNow) {   // the partial sums read in parallel, added in tile order by one thread
            const double* P = ta.part + td.pOff + (size_t)par * td.nTiles * 4;
            double* sP = (double*)sF;   // the load buffer is free between the voxel and link passes
            for (int j = tid; j < td.nTiles; j += BS) {
                const double a0 = xget(P + 4 * j), a1 = xget(P + 4 * j + 1), a2 = xget(P + 4 * j + 2), a3 = xget(P + 4 * j + 3);
                sP[4 * j] = a0; sP[4 * j + 1] = a1; sP[4 * j + 2] = a2; sP[4 * j + 3] = a3;
            }
            __syncthreads();
            if (tid == 0) {   // the same totals in every tile
                for (int q = 0; q < 4; q++) sRed[q][0] = 0;
                for (int j = 0; j < td.nTiles; j++)
                    for (int q = 0; q < 4; q++) sRed[q][0] += sP[4 * j + q];
            }
        }
        if (loadHalo) {
            const real* X = (const real*)ta.xch + td.xOff + (size_t)par * td.xPar;
#pragma unroll 1
            for (int v = nvOwn + tid; v < nvAll; v += BS) {
                // every load issued before the first LDS write: one memory latency, not eleven
                const real* r = X + (size_t)sXv[v] * VX3_XCH_STRIDE;
                real q[VX3_XCH_STRIDE - 1];
#pragma unroll
                for (int f = 0; f < VX3_XCH_STRIDE - 1; f++) q[f] = xget(r + f);
                sPx[v] = q[0]; sPy[v] = q[1]; sPz[v] = q[2];
                sOw[v] = q[3]; sOx[v] = q[4]; sOy[v] = q[5]; sOz[v] = q[6];
                sTf[v] = (float)q[7]; sPsx[v] = (float)q[8]; sPsy[v] = (float)q[9]; sPsz[v] = (float)q[10];
            }
            __syncthreads();
        }
        return 0;
    };
    for (int it = 0; it < nSteps && !colOvf; it++) {
        // ---- contact lists rebuilt when stale or at the fixed period (COL)
        if (COL && ((it == 0 && colFirst) || sStale[it & 1] != 0 || (scg.colInterval > 0 && steps % scg.colInterval == 0))) {
            int* cnt = (int*)sF;                        // [VMAX] voxels per bucket
            int* bst = cnt + VMAX;                      // [VMAX] first sorted voxel of each bucket
            unsigned short* cvox = (unsigned short*)(bst + VMAX);   // [VMAX] surface voxels by bucket
            constexpr unsigned mask = VMAX - 1;
            const double cellE = (double)scg.colCell;
            for (int i = tid; i < VMAX; i += BS) cnt[i] = 0;
            __syncthreads();
            int bk[VPT], rk[VPT];
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                bk[k] = -1;
                rk[k] = 0;
                const int v = tid + k * BS;
                if (vmat[k] < 0 || (vf[k] & VF_INTERIOR)) continue;
                bk[k] = (int)(colHash((int)floor((double)sPx[v] / cellE), (int)floor((double)sPy[v] / cellE),
                                      (int)floor((double)sPz[v] / cellE)) & mask);
                rk[k] = atomicAdd(&cnt[bk[k]], 1);
                lw[k][0] = (float)sPx[v]; lw[k][1] = (float)sPy[v]; lw[k][2] = (float)sPz[v];   // lastColWatchPosition
            }
            __syncthreads();
            int loc = 0;   // thread tid owns buckets [tid*VPT, tid*VPT + VPT)
#pragma unroll
            for (int k = 0; k < VPT; k++) loc += cnt[tid * VPT + k];
            int tot;
            int off = resBlockScan<BS>(loc, sScan, &tot);
#pragma unroll
            for (int k = 0; k < VPT; k++) { bst[tid * VPT + k] = off; off += cnt[tid * VPT + k]; }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < VPT; k++)
                if (bk[k] >= 0) cvox[bst[bk[k]] + rk[k]] = (unsigned short)(tid + k * BS);
            __syncthreads();
            // every watched partner of surface voxel v, from the voxels of its 27 neighbouring cells
            auto forPartners = [&](int v, auto f) {
                const int cx = (int)floor((double)sPx[v] / cellE), cy = (int)floor((double)sPy[v] / cellE),
                          cz = (int)floor((double)sPz[v] / cellE);
                int hs[27];
#pragma unroll
                for (int q = 0; q < 27; q++) hs[q] = (int)(colHash(cx + q % 3 - 1, cy + (q / 3) % 3 - 1, cz + q / 9 - 1) & mask);
                for (int q = 0; q < 27; q++) {
                    const int h = hs[q];
                    bool seen = false;   // a bucket shared by several of the 27 cells is read once
                    for (int j = 0; j < q; j++) seen |= hs[j] == h;
                    if (seen) continue;
                    const int n = cnt[h], b0 = bst[h];
                    for (int i = 0; i < n; i++) {
                        const int u = cvox[b0 + i];
                        if (u != v && resWatched(b, scg, sPx, sPy, sPz, v, u)) f(u);
                    }
                }
            };
            int pc[VPT], mine = 0;
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                pc[k] = 0;
                if (bk[k] >= 0) forPartners(tid + k * BS, [&](int) { pc[k]++; });
                mine += pc[k];
            }
            int total;
            int o = resBlockScan<BS>(mine, sScan, &total);
            if (total > VX3_RES_PAIRS || total > scg.nbCap) { colOvf = true; break; }   // uniform: the block's total
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                nbS[k] = o;
                nbC[k] = pc[k];
                o += pc[k];
                if (!pc[k]) continue;
                unsigned short* out = sNb + nbS[k];
                int n = 0;
                forPartners(tid + k * BS, [&](int u) { out[n++] = (unsigned short)u; });
                for (int i = 1; i < n; i++) {   // ascending partner order
                    const unsigned short x = out[i];
                    int j = i - 1;
                    while (j >= 0 && out[j] > x) { out[j + 1] = out[j]; j--; }
                    out[j + 1] = x;
                }
            }
            __syncthreads();
            rebuilds++;
        }
        // keep per-step LDS address arithmetic inside the loop (register pressure)
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int k = 0; k < LPA; k++) { asm volatile("" : "+v"(lEnds[a][k])); asm volatile("" : "+v"(lMats[a][k])); }
#pragma unroll
        for (int k = 0; k < VPT; k++) { asm volatile("" : "+v"(vl[k][0])); asm volatile("" : "+v"(vl[k][1])); asm volatile("" : "+v"(vl[k][2])); asm volatile("" : "+v"(vmat[k])); }
        // temperatures of this step's distinct phases (read in the voxel pass, after the link barriers)
        if (vary && tid < nph) sPhT[tid] = actuationTemp(tempAmp, tempPeriod, sPhase[tid], t + dt, expansion);
        V3<real> fs[VPT], ms[VPT];
#pragma unroll
        for (int k = 0; k < VPT; k++) { fs[k] = V3<real>(0, 0, 0); ms[k] = V3<real>(0, 0, 0); }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            // ---- link pass of axis a.  TILED with the last step's barrier pending: the
            // axis-0 links between owned voxels first, then the wait and the halo, then the
            // axis-0 links with a halo end
            auto linkAt = [&](int k) {
                const int slot = tid + k * BS;
                const int vn = lEnds[a][k] & 0xFFFF, vq = lEnds[a][k] >> 16;
                const LinkMatDev& mat = sLM[lMats[a][k] & 0xFF];
                const VoxMatDev& mn = sVM[(lMats[a][k] >> 8) & 0xFF];
                const VoxMatDev& mp = sVM[lMats[a][k] >> 16];
                if (mn.fixed && mp.fixed) return;   // its slots are never read: both ends are fixed
                const real restLength = (real)0.5 * (nom * (real)sTf[vn] + nom * (real)sTf[vq]);
                float area, tss;
                if (mat.nu != 0.0f) {
                    float an, sn, ap, sp;
                    // CVX_Voxel::transverseArea / transverseStrainSum from the pose cache
                    auto tr = [&](int v, bool nuZero, float& ar, float& su) {
                        if (nuZero) { ar = size * size; su = 0.0f; return; }
                        const float sx = sPsx[v], sy = sPsy[v], sz = sPsz[v];
                        const real s2 = (real)(size * size);
                        if (a == 0) { ar = (float)(s2 * ((real)1 + (real)sy) * ((real)1 + (real)sz)); su = sy + sz; }
                        else if (a == 1) { ar = (float)(s2 * ((real)1 + (real)sx) * ((real)1 + (real)sz)); su = sx + sz; }
                        else { ar = (float)(s2 * ((real)1 + (real)sx) * ((real)1 + (real)sy)); su = sx + sy; }
                    };
                    tr(vn, mn.nu == 0.0f, an, sn);
                    tr(vq, mp.nu == 0.0f, ap, sp);
                    area = 0.5f * (an + ap);
                    tss = 0.5f * (sn + sp);
                } else {
                    area = 0.5f * (size * size + size * size);
                    tss = 0.0f;
                }
                const float dmN = sDm[(lMats[a][k] >> 8) & 0xFF];
                const float dmP = sDm[lMats[a][k] >> 16];
                const V3<real> pN(sPx[vn], sPy[vn], sPz[vn]), pP(sPx[vq], sPy[vq], sPz[vq]);
                const Q4<real> oN(sOw[vn], sOx[vn], sOy[vn], sOz[vn]), oP(sOw[vq], sOx[vq], sOy[vq], sOz[vq]);
                V3<real> fN, mN, fP, mP;
                linkForces<real>(mat, a, pN, oN, pP, oP, restLength, area, tss, mat.nu != 0.0f, dmN, dmP, ls[a][k], fN, mN, fP, mP);
                sF[0 * LAX + slot] = fN.x; sF[1 * LAX + slot] = fN.y; sF[2 * LAX + slot] = fN.z;
                sF[3 * LAX + slot] = mN.x; sF[4 * LAX + slot] = mN.y; sF[5 * LAX + slot] = mN.z;
                sF[6 * LAX + slot] = fP.x; sF[7 * LAX + slot] = fP.y; sF[8 * LAX + slot] = fP.z;
                sF[9 * LAX + slot] = mP.x; sF[10 * LAX + slot] = mP.y; sF[11 * LAX + slot] = mP.z;
                if (poisson) {
                    // CVX_Link::axialStrain(positiveEnd) (VX_Link.cpp:115-118)
                    const float r = mat.strainRatio;
                    sEs[(2 * a) * LAX + slot] = 2.0f * ls[a][k].strain / (1.0f + r);
                    sEs[(2 * a + 1) * LAX + slot] = 2.0f * ls[a][k].strain * r / (1.0f + r);
                }
                if (ls[a][k].strain > 100) sDiv = 1;
            };
            if (TILED && a == 0 && pend) {
#pragma unroll
                for (int k = 0; k < LPA; k++)
                    if (lEnds[a][k] >= 0 && (lEnds[a][k] & 0xFFFF) < nvOwn && (lEnds[a][k] >> 16) < nvOwn) linkAt(k);
                if (finishBarrier(true, false)) return;
#pragma unroll
                for (int k = 0; k < LPA; k++)
                    if (lEnds[a][k] >= 0 && ((lEnds[a][k] & 0xFFFF) >= nvOwn || (lEnds[a][k] >> 16) >= nvOwn)) linkAt(k);
            } else {
#pragma unroll
                for (int k = 0; k < LPA; k++)
                    if (lEnds[a][k] >= 0) linkAt(k);
            }
            __syncthreads();
            if (COL && a == 0 && tid == 0) sStale[it & 1] = 0;   // every thread has read it
            // ---- the voxels add this axis' two directions: 2a (negative end) then 2a+1 (positive end)
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                if (vmat[k] < 0) continue;
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int slot = e ? ((vl[k][a] >> 16) & 0xFFFF) : (vl[k][a] & 0xFFFF);
                    if (slot == 0xFFFF) continue;
                    const int c = 6 * e;
                    fs[k].add(V3<real>(sF[(c + 0) * LAX + slot], sF[(c + 1) * LAX + slot], sF[(c + 2) * LAX + slot]));
                    ms[k].add(V3<real>(sF[(c + 3) * LAX + slot], sF[(c + 4) * LAX + slot], sF[(c + 5) * LAX + slot]));
                }
            }
            if (a < 2) __syncthreads();   // the next axis overwrites the load buffer
        }
        if (COL) {   // the load buffer becomes the step-start snapshot the contacts read
            __syncthreads();
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                if (vmat[k] < 0) continue;
                const int v = tid + k * BS;
                sF[v] = sPx[v]; sF[VMAX + v] = sPy[v]; sF[2 * VMAX + v] = sPz[v];
                sF[3 * VMAX + v] = vm[k].x; sF[4 * VMAX + v] = vm[k].y; sF[5 * VMAX + v] = vm[k].z;
            }
            __syncthreads();
        }
        // written before the last link-pass barrier; a tile learns of its simulation's
        // divergence only at the step's barrier and stops at that step on a second run
        const bool diverged = multi ? it == divAt : sDiv != 0;
        steps++;
        const bool sample = !diverged && cycleStep > 0 && steps % cycleStep == 0;
        double psx = 0, psy = 0, psz = 0, psm = 0;
        const double tn = t + dt;
        float newTf[VPT];
#pragma unroll
        for (int k = 0; k < VPT; k++) newTf[k] = -1.0f;
        // ---- voxel pass
        if (!diverged) {
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                if (vmat[k] < 0) continue;
                const int v = tid + k * BS;
                const VoxMatDev& m = sVM[vmat[k]];
                V3<real> vp(sPx[v], sPy[v], sPz[v]);
                if (!m.fixed) {
                    const ExtDev* ext = vext[k] >= 0 ? &b.extTab[vext[k]] : nullptr;
                    int ix = 0, iy = 0, iz = 0;
                    if (ext) { ix = ext->ix; iy = ext->iy; iz = ext->iz; }
                    Q4<real> vo(sOw[v], sOx[v], sOy[v], sOz[v]);
                    if (COL && nbC[k] > 0) {
                        const ResContacts con{&b, sF, sTf, sNb, VMAX, v, v0, nbS[k], nbC[k], nom};
                        voxelStep<real>(m, ext, dtf, nom, ix, iy, iz, sTf[v], fs[k], ms[k], con, vp, vm[k], vo, va[k], vf[k]);
                    } else {
                        voxelStep<real>(m, ext, dtf, nom, ix, iy, iz, sTf[v], fs[k], ms[k], NoContacts(), vp, vm[k], vo, va[k], vf[k]);
                    }
                    sPx[v] = vp.x; sPy[v] = vp.y; sPz[v] = vp.z;
                    sOw[v] = vo.w; sOx[v] = vo.x; sOy[v] = vo.y; sOz[v] = vo.z;
                }
                if (COL && !(vf[k] & VF_INTERIOR)) {   // CVoxelyze::updateCollisions' staleness rule, for the next step
                    const real dx = vp.x - (real)lw[k][0], dy = vp.y - (real)lw[k][1], dz = vp.z - (real)lw[k][2];
                    if (dx * dx + dy * dy + dz * dz > (real)scg.colRecalcSq) sStale[(it + 1) & 1] = 1;
                }
                if (vary && !m.fixed && !(m.thermalOnAfter > tn)) {
                    // with contacts, every voxel reads this step's factors first (written below)
                    if (COL) newTf[k] = 1 + sPhT[vpi[k]] * m.alphaCTE;
                    else sTf[v] = 1 + sPhT[vpi[k]] * m.alphaCTE;
                }
                if (poisson) {
                    // CVX_Voxel::strain(true) for the next step from this step's end strains:
                    // per axis the sum of the present ends in direction order, halved when
                    // both are present; the axes not in tension take the Poisson value
                    float st[3];
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        const int s0 = vl[k][i] & 0xFFFF, s1 = (vl[k][i] >> 16) & 0xFFFF;
                        const float x0 = sEs[(2 * i) * LAX + (s0 != 0xFFFF ? s0 : 0)];
                        const float x1 = sEs[(2 * i + 1) * LAX + (s1 != 0xFFFF ? s1 : 0)];
                        float sum = s0 != 0xFFFF ? 0.0f + x0 : 0.0f;
                        if (s1 != 0xFFFF) sum += x1;
                        st[i] = (vpm[k] & (1 << i)) ? sum * 0.5f : sum;
                    }
                    if ((vpm[k] & 56) != 56) {
                        float add = 0;
#pragma unroll
                        for (int i = 0; i < 3; i++) if (vpm[k] & (8 << i)) add += st[i];
                        const float value = poissonValue(1.0f + add, m.nu);
#pragma unroll
                        for (int i = 0; i < 3; i++) if (!(vpm[k] & (8 << i))) st[i] = value;
                    }
                    sPsx[v] = st[0]; sPsy[v] = st[1]; sPsz[v] = st[2];
                }
                if (sample && m.isMeasured) {
                    const double mass = m.mass;
                    psx += (double)vp.x * mass; psy += (double)vp.y * mass; psz += (double)vp.z * mass;
                    psm += mass;
                }
                if (TILED && sXv[v] >= 0) {
                    // a voxel another tile reads: its record out as soon as it is final (such
                    // voxels take the lowest slots, so the stores drain while the rest integrate)
                    real q[VX3_XCH_STRIDE - 1] = {sPx[v], sPy[v], sPz[v], sOw[v], sOx[v], sOy[v], sOz[v],
                                                  (real)sTf[v], (real)sPsx[v], (real)sPsy[v], (real)sPsz[v]};
                    real* r = (real*)ta.xch + td.xOff + (size_t)(gen & 1) * td.xPar + (size_t)sXv[v] * VX3_XCH_STRIDE;
#pragma unroll
                    for (int f = 0; f < VX3_XCH_STRIDE - 1; f++) xput(r + f, q[f]);
                }
            }
        }
        if (COL && !diverged) {
            __syncthreads();   // every contact has read this step's thermal factors
#pragma unroll
            for (int k = 0; k < VPT; k++)
                if (newTf[k] >= 0.0f) sTf[tid + k * BS] = newTf[k];
        }
        if (sample) {
            // FP64 center of mass of the measured voxels (VX3_VoxelyzeKernel.cu:477-493)
            for (int o = 32; o > 0; o >>= 1) {
                psx += __shfl_xor(psx, o, 64); psy += __shfl_xor(psy, o, 64);
                psz += __shfl_xor(psz, o, 64); psm += __shfl_xor(psm, o, 64);
            }
            if ((tid & 63) == 0) { sRed[0][tid >> 6] = psx; sRed[1][tid >> 6] = psy; sRed[2][tid >> 6] = psz; sRed[3][tid >> 6] = psm; }
        }
        if (multi && !diverged) {
            // ---- arrival at the simulation's barrier (TILED): the records are out; the wait
            // comes in the next step's link pass, or now on a sampling step (the centre of
            // mass needs every tile's partial sum)
            __builtin_amdgcn_s_waitcnt(0);   // this thread's exchange stores are complete
            __syncthreads();
            if (tid == 0) {
                if (sample) {
                    double* P = ta.part + td.pOff + (size_t)(gen & 1) * td.nTiles * 4;
                    double tot[4] = {0, 0, 0, 0};
                    for (int w = 0; w < BS / 64; w++)
                        for (int q = 0; q < 4; q++) tot[q] += sRed[q][w];
                    for (int q = 0; q < 4; q++) xput(P + 4 * td.tile + q, tot[q]);
                }
                if (sDiv) __hip_atomic_fetch_max(ta.sync + td.cOff, it + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                tileArrive(ta.sync + ta.flag0 + blockIdx.x, gen, ta.repair);
            }
            pend = true;
            if (sample && finishBarrier(true, true)) return;
        } else if (sample) __syncthreads();
        if (sample) {
            angleSamples++;
            if (tid == 0) {
                double tot[4] = {0, 0, 0, 0};
                if (multi) {
                    for (int q = 0; q < 4; q++) tot[q] = sRed[q][0];
                } else {
                    for (int w = 0; w < BS / 64; w++)
                        for (int q = 0; q < 4; q++) tot[q] += sRed[q][w];
                }
                double* A = sComH0; double* B = sComH1; double* C = sCom;
                for (int q = 0; q < 3; q++) { A[q] = B[q]; B[q] = C[q]; C[q] = tot[3] == 0 ? 0.0 : (1.0 / tot[3]) * tot[q]; }
                const bool bc = B[0] == C[0] && B[1] == C[1] && B[2] == C[2];
                const bool ab = A[0] == B[0] && A[1] == B[1] && A[2] == B[2];
                if (bc || ab || angleSamples < 3) sAngle = 0;
                else {
                    double d1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, d2[3] = {C[0] - B[0], C[1] - B[1], C[2] - B[2]};
                    double dot = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
                    double l1 = sqrt((A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]));
                    double l2 = sqrt((B[0] - C[0]) * (B[0] - C[0]) + (B[1] - C[1]) * (B[1] - C[1]) + (B[2] - C[2]) * (B[2] - C[2]));
                    sAngle = acos(dot / (l1 * l2));
                }
            }
        }
        // ---- per-simulation control (uniform across the workgroup)
        if (diverged) { state = 2; break; }
        t = tn;
        bool stop = maxSteps > 0 && steps >= maxSteps;
        if (!stop && stopLen > 0 && tid == 0) sRun = mathTreeEval(b.tokens + stopOff, stopLen, sCom[0], sCom[1], sCom[2], 0, t, sAngle, 0, 0, nv, sQ) > 0 ? 0 : 1;
        if (it == 0 && prevDt0 != dtf && tid < nvm) sDm[tid] = 2 * sVM[tid].sqrtMass * sVM[tid].zetaInternal / (sVM[tid].fixed ? 0.0f : dtf);
        __syncthreads();   // next step's link pass reads the new poses
        if (stop || !sRun) {
            if (TILED && pend && finishBarrier(false, false)) return;
            state = 1;
            break;
        }
    }
    if (TILED && pend && finishBarrier(false, false)) return;   // the launch's last step

    // ---- write back in the stream layout
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int k = 0; k < LPA; k++) {
            if (lEnds[a][k] < 0) continue;
            if (TILED && !(ta.link[td.lOff + lIdx[a][k]] & 0x80000000)) continue;   // the other tile writes it
            const int g = glink(lIdx[a][k]);
            b.q2x[g] = ls[a][k].pos2.x; b.q2y[g] = ls[a][k].pos2.y; b.q2z[g] = ls[a][k].pos2.z;
            b.a1x[g] = ls[a][k].a1v.x; b.a1y[g] = ls[a][k].a1v.y; b.a1z[g] = ls[a][k].a1v.z;
            b.a2x[g] = ls[a][k].a2v.x; b.a2y[g] = ls[a][k].a2v.y; b.a2z[g] = ls[a][k].a2v.z;
            b.strain[g] = ls[a][k].strain;
            b.maxStrain[g] = ls[a][k].maxStrain;
            b.strainOffset[g] = ls[a][k].strainOffset;
            b.lflags[g] = ls[a][k].flags;
            if (poisson) {
                const int slot = tid + k * BS;
                const int vn = lEnds[a][k] & 0xFFFF, vq = lEnds[a][k] >> 16;
                const int gn = TILED ? b.lneg[g] : v0 + vn, gq = TILED ? b.lpos[g] : v0 + vq;
                b.es[(2 * a) * VC + gn] = sEs[(2 * a) * LAX + slot];
                b.es[(2 * a + 1) * VC + gq] = sEs[(2 * a + 1) * LAX + slot];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        if (vmat[k] < 0) continue;
        const int v = tid + k * BS, g = gvox(v);
        b.px[g] = sPx[v]; b.py[g] = sPy[v]; b.pz[g] = sPz[v];
        b.mx[g] = vm[k].x; b.my[g] = vm[k].y; b.mz[g] = vm[k].z;
        b.ax[g] = va[k].x; b.ay[g] = va[k].y; b.az[g] = va[k].z;
        b.ow[g] = sOw[v]; b.ox[g] = sOx[v]; b.oy[g] = sOy[v]; b.oz[g] = sOz[v];
        b.vflags[g] = vf[k];
        b.tfac[g] = sTf[v];
        b.psx[g] = sPsx[v]; b.psy[g] = sPsy[v]; b.psz[g] = sPsz[v];
    }
    if (COL && !colOvf) {   // contact lists and last-rebuild positions in the stream layout
        const int nb0 = scg.nb0;
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            if (vmat[k] < 0 || (vf[k] & VF_INTERIOR)) continue;
            const int g = v0 + tid + k * BS;
            b.lastWatch[3 * g] = lw[k][0]; b.lastWatch[3 * g + 1] = lw[k][1]; b.lastWatch[3 * g + 2] = lw[k][2];
            b.nbStart[g] = nb0 + nbS[k];
            b.nbCount[g] = nbC[k];
            for (int i = 0; i < nbC[k]; i++) b.nbList[nb0 + nbS[k] + i] = v0 + sNb[nbS[k] + i];
        }
    }
    if (tid == 0 && (!TILED || td.tile == 0)) {   // every tile holds the same control values
        if (COL) {
            int f = scg.flags;
            if (rebuilds > 0 && !colOvf) f &= ~SC_REBUILD;
            if (colOvf) f |= SC_COL_GROW | SC_REBUILD;   // the host moves it to the per-step kernels
            scg.flags = f;
            scg.rebuilds += rebuilds;
        }
        if (steps != scg.steps) scg.prevDt = dtf;
        scg.t = t;
        scg.steps = steps;
        scg.angleSamples = angleSamples;
        for (int q = 0; q < 3; q++) { scg.com[q] = sCom[q]; scg.comH0[q] = sComH0[q]; scg.comH1[q] = sComH1[q]; }
        scg.recentAngle = sAngle;
        scg.state = state;
    }
}

}  // namespace VX_NS
}  // namespace vx3
