// vx3h_engine.cpp -- the batch engine behind the C ABI (include/vx3h.h).
//
// A batch owns one HIP stream and structure-of-arrays device buffers for many
// simulations.  Simulations occupy contiguous voxel and link ranges handed out
// by first-fit free lists, so finished simulations can be retired and new ones
// admitted between launches (continuous batching).  Every kernel covers the
// high-water mark of the ranges and skips elements whose simulation is not
// running.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdarg>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vx3h.h"
#include "vx3_launch.h"
#include "vx3h_build.h"
#include "vx3h_internal.h"
#include "vx3h_tile.h"

namespace vx3 {
thread_local std::string g_err;
}
using vx3::g_err;

#define FAILV(code, msg)        \
    do {                        \
        g_err = (msg);          \
        return (code);          \
    } while (0)
#define HIPCHK(x)                                                                 \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            g_err = std::string(#x) + " failed: " + hipGetErrorString(e_);        \
            return VX3H_EDEVICE;                                                  \
        }                                                                         \
    } while (0)

extern "C" const char* vx3h_last_error(void) { return g_err.c_str(); }

extern "C" vx3h_status vx3h_device_count(int* count) {
    if (!count) FAILV(VX3H_EINVAL, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return VX3H_OK;
}

namespace {

using namespace vx3;

// One device array family: `per` values of `elem` bytes per element, stored
// either interleaved (index*per + k) or strided by capacity (k*cap + index).
struct ArrSpec {
    void** slot;
    size_t elem;
    int per;
    bool strided;
};

struct Range { int start, len; };

struct RangeAlloc {
    std::map<int, int> freeRanges;   // start -> length
    int hi = 0;
    int alloc(int n, int cap, bool& grow) {
        grow = false;
        for (auto it = freeRanges.begin(); it != freeRanges.end(); ++it) {
            if (it->second >= n) {
                int s = it->first, len = it->second;
                freeRanges.erase(it);
                if (len > n) freeRanges[s + n] = len - n;
                return s;
            }
        }
        int s = hi;
        hi += n;
        if (hi > cap) grow = true;
        return s;
    }
    void release(int s, int n) {
        if (n <= 0) return;
        freeRanges[s] = n;
        auto it = freeRanges.find(s);
        auto nx = std::next(it);
        if (nx != freeRanges.end() && it->first + it->second == nx->first) { it->second += nx->second; freeRanges.erase(nx); }
        if (it != freeRanges.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; freeRanges.erase(it); it = pv; }
        }
        if (it->first + it->second == hi) { hi = it->first; freeRanges.erase(it); }
    }
};

struct SlotHost {
    bool used = false;
    int v0 = 0, nv = 0, l0 = 0, nl = 0;
    int nlCap = 0;               // link range size: nl, or room for the links attach makes
    int nlSeen = 0;              // links known to the host shadows below (nl + attached ones read back)
    int shape = -1;              // resident kernel shape, -1 = stream path (or tiled)
    int tiles = 0;               // > 0: runs on the tiled resident kernel as this many workgroups
    TilePlan tplan;              // its tiles (sim-local; the tables themselves live on the device)
    int tv0 = 0, tvN = 0, tl0 = 0, tlN = 0, tx0 = 0, txN = 0, tp0 = 0;   // its ranges in the tile pools
    bool oneTile = false;        // a one-workgroup simulation with one-tile tables in the pools
    int axisLinks = 0;           // links on its busiest axis
    std::vector<int> ijk;
    std::vector<int> matid;
    std::vector<int> measured;   // per voxel
    std::vector<float> mass;     // per voxel
    double voxSize = 0;
    double com0[3] = {0, 0, 0};
    std::vector<vx3h_token> fit;
    size_t bytes = 0;
    // collision pools: neighbour lists, hash buckets, nearby lists
    int nb0 = 0, nbN = 0, h0 = 0, hN = 0, near0 = 0, nearN = 0, c0 = 0;   // hN in buckets (whole blocks)
    int tg0 = 0, nTg = 0;        // close pairs: range of the sim's target voxels in tgtList
    int at0 = 0, ap0 = 0, apN = 0;   // attach scratch ranges (attCand: 2*nv; attPair: apN)
    // history records: topology and link constants of the sim, palette in order
    std::vector<int> vlinks;     // 6 per voxel, sim-local link index or -1
    std::vector<char> surface;   // per voxel: not all six links present
    std::vector<int> lneg, lpos; // per link, sim-local voxel indices
    std::vector<float> lratio, lfail;   // per link: strainRatio, epsilonFail (-1 = none)
    std::vector<std::array<int, 5>> palette;   // matid, r, g, b, a
};

// Device arrays sharing one index space of first-fit ranges (collision pools).
struct DevPool {
    std::vector<std::pair<void**, size_t>> arrs;   // slot, bytes per element
    int cap = 0;
    int minCap = 4096;                             // elements allocated at first use
    RangeAlloc ra;
};

template <typename K> struct Dedup {
    std::vector<K> items;
    std::map<std::string, int> index;
    int add(const K& k) {
        std::string key((const char*)&k, sizeof(K));
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        int i = (int)items.size();
        items.push_back(k);
        index[key] = i;
        return i;
    }
};

}  // namespace

struct vx3h_batch {
    int device = 0;
    vx3h_batch_cfg cfg{};
    hipStream_t stream = nullptr;
    size_t realSize = 4;
    DevBatch<float> db{};   // pointer layout shared with DevBatch<double>
    std::vector<ArrSpec> vArrs, lArrs;
    RangeAlloc va, la;
    std::vector<SlotHost> slots;
    int simCap = 0;
    Dedup<VoxMatDev> vmats;
    Dedup<LinkMatDev> lmats;
    std::vector<ExtDev> exts;
    std::vector<TokenDev> tokens;
    std::vector<int> simMats;                    // per sim: global material table indices it uses
    void* dVmat = nullptr; size_t capVmat = 0;
    void* dLmat = nullptr; size_t capLmat = 0;
    void* dExt = nullptr; size_t capExt = 0;
    void* dTok = nullptr; size_t capTok = 0;
    void* dSimMats = nullptr; size_t capSimMats = 0;
    DevPool nbPool, hashPool, nearPool, cellPool, tgtPool, attPool, pairPool;
    bool anyColl = false;
    bool anyExtras = false;             // some simulation has cilia or a force field
    bool anyDetach = false;             // some simulation detaches failed links
    bool anyAttach = false;             // some simulation attaches sticky voxels
    bool anySignals = false;            // some simulation has signals
    bool anyTopoCol = false;            // some simulation has collisions and attach, detach or voxel removal
    int dtLo = 1 << 30, dtHi = -1;
    size_t bytesUsed = 0, budget = 0;
    // resident path: kernel shapes and, per shape, the sims it runs
    std::vector<ResidentShapeInfo> shapes;
    std::vector<int> shapePref;                  // shape indices in preference order
    std::vector<std::vector<int>> resLists;
    std::vector<int*> dResLists;
    std::vector<size_t> capRes;
    bool resDirty = false;
    int nStreamSims = 0;
    // tiled resident path: simulations too large for one workgroup run as several tiles,
    // packed into cooperative launches of at most tileCap workgroups (whole simulations),
    // one after another on tileStream beside the other paths
    int tiledShape = -1;                         // shape of the tiled kernel: the largest voxel capacity
    int tileCap = 0;                             // workgroups of it the device holds at once (0: no tiling)
    long long spinTicks = 0;                     // a tile's longest barrier wait, in wall-clock ticks
    std::vector<int> tiledSims;                  // slots, in admission order
    bool tileDirty = false;
    struct TileGroup { int first, count, sync0, nSync; };
    std::vector<TileGroup> tileGroups;
    TileDesc* dTiles = nullptr; size_t capTiles = 0;
    int* dSync = nullptr; size_t capSync = 0;
    // per simulation, uploaded once at admission: tile slots (voxel, exchange record), tile
    // links (link, end slots), exchange records (two step parities), partial sums per tile
    DevPool tileSlotPool, tileLinkPool, tileXchPool, tilePartPool;
    int* dTileVox = nullptr; int* dTileXidx = nullptr;
    int* dTileLink = nullptr; int* dTileEnds = nullptr;
    void* dXch = nullptr;
    double* dPart = nullptr;
    hipStream_t tileStream = nullptr;
    hipEvent_t tileJoin = nullptr;
    bool tilePending = false;
    // while tiled simulations run, the one-workgroup simulations without collisions run in
    // the same cooperative launches as one-tile entries (a cooperative launch can only
    // start once every CU it needs is free: beside them it would wait for theirs)
    bool absorbed = false;
    // one stream per resident shape: each shape's launches run beside the stream-path
    // steps of the main stream (disjoint simulations) and beside the other shapes;
    // fork/join events order them with the main stream around every chunk of steps
    std::vector<hipStream_t> resStreams;
    std::vector<hipEvent_t> joinEv;
    std::vector<bool> joinPending;
    hipEvent_t forkEv = nullptr;
    // graph of `graphSteps` steps, valid for one layout version
    hipGraphExec_t graphExec = nullptr;
    hipGraph_t graph = nullptr;
    int graphSteps = 0;
    long layout = 0, graphLayout = -1;
    // launch timing (vx3h_batch_set_timing)
    bool timing = false;
    std::vector<hipEvent_t> evPool;              // free events
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evUsed;
    long long timedSteps = 0;
    int tableGrowths = 0;                        // collision neighbour tables enlarged so far
    // A simulation that stood still for a collision table after its link pass resumes at
    // its next step from the end loads that link pass left, in the buffer the stepping
    // entry point of that step used: the batch's own (false) or the no-allocation
    // entry point's static array (true).
    bool loadsInScratch = false;
    size_t loadsFs = 0;   // slot stride of that buffer at that step
    int loadsNv = 0;      // voxels [0, loadsNv) it holds
    bool resumePending = false;
    // host-to-device uploads of admissions: staged in pinned host memory and copied on the
    // main stream, so an admission waits for the device once (at the next admission's
    // start) rather than once per array; a region is reused only after the stream has run
    char* stage = nullptr;
    size_t stageCap = 0, stageUsed = 0;
};

namespace {

size_t voxelBytes(size_t rs) { return rs * (13 + 36 + 6 + 6) + 4 * 5 + 4 * 3 + 4 * 4 + 4 * 6 + 4 + 8 + 8 * 8 + 8 * 3 + 8 + 8 * 8 + 8 * 5 + 4 * 6 + 4 * 2; }
size_t linkBytes(size_t rs) { return rs * 9 + 4 * 6 + 4 * 3; }

void initSpecs(vx3h_batch* b) {
    DevBatch<float>& d = b->db;
    const size_t R = b->realSize;
    auto V = [&](void* p, size_t e, int per, bool st) { b->vArrs.push_back({(void**)p, e, per, st}); };
    auto L = [&](void* p, size_t e, int per, bool st) { b->lArrs.push_back({(void**)p, e, per, st}); };
    V(&d.px, R, 1, false); V(&d.py, R, 1, false); V(&d.pz, R, 1, false);
    V(&d.mx, R, 1, false); V(&d.my, R, 1, false); V(&d.mz, R, 1, false);
    V(&d.ow, R, 1, false); V(&d.ox, R, 1, false); V(&d.oy, R, 1, false); V(&d.oz, R, 1, false);
    V(&d.ax, R, 1, false); V(&d.ay, R, 1, false); V(&d.az, R, 1, false);
    V(&d.vflags, 4, 1, false); V(&d.tfac, 4, 1, false); V(&d.phase, 8, 1, false);
    V(&d.vmat, 4, 1, false); V(&d.vsim, 4, 1, false); V(&d.vext, 4, 1, false); V(&d.vijk, 4, 3, false);
    V(&d.psx, 4, 1, false); V(&d.psy, 4, 1, false); V(&d.psz, 4, 1, false);
    V(&d.F, R, 36, true); V(&d.es, 4, 6, true);
    V(&d.lastWatch, 4, 3, false); V(&d.nbStart, 4, 1, false); V(&d.nbCount, 4, 1, false);
    V(&d.vbk, 4, 1, false); V(&d.vrank, 4, 1, false);
    V(&d.cpx, R, 1, false); V(&d.cpy, R, 1, false); V(&d.cpz, R, 1, false);
    V(&d.cmx, R, 1, false); V(&d.cmy, R, 1, false); V(&d.cmz, R, 1, false); V(&d.ctf, 4, 1, false);
    V(&d.nearStart, 4, 1, false); V(&d.nearCount, 4, 1, false); V(&d.nearMask, 4, VX3_NEAR_WORDS, true);
    V(&d.cil, 8, 8, true); V(&d.xf, R, 6, true); V(&d.ip, 8, 3, true); V(&d.rt, 8, 1, false);
    V(&d.sg, 8, 8, true); V(&d.sp, 8, 5, true); V(&d.nbr, 4, 6, true);
    b->nbPool.arrs = {{(void**)&d.nbList, 4}};
    // the hash pool is allocated in blocks of VX3_COL_HBLOCK buckets
    b->hashPool.arrs = {{(void**)&d.hashCnt, 4 * VX3_COL_HBLOCK}, {(void**)&d.hashStart, 4 * VX3_COL_HBLOCK},
                        {(void**)&d.hashSim, 4 * VX3_COL_HBLOCK}, {(void**)&d.hashBlk, 4}, {(void**)&d.hashBlkOff, 4}};
    b->cellPool.arrs = {{(void**)&d.cellVox, 4}};
    b->hashPool.minCap = 16;
    b->tgtPool.arrs = {{(void**)&d.tgtList, 4}};
    b->attPool.arrs = {{(void**)&d.attCand, 4}};
    b->pairPool.arrs = {{(void**)&d.attPair, 8}};
    b->nearPool.arrs = {{(void**)&d.nearList, 4}};
    b->tileSlotPool.arrs = {{(void**)&b->dTileVox, 4}, {(void**)&b->dTileXidx, 4}};
    b->tileLinkPool.arrs = {{(void**)&b->dTileLink, 4}, {(void**)&b->dTileEnds, 4}};
    b->tileXchPool.arrs = {{(void**)&b->dXch, 2 * VX3_XCH_STRIDE * R}};   // one record, both parities
    b->tilePartPool.arrs = {{(void**)&b->dPart, 8 * sizeof(double)}};    // one tile, both parities
    b->tileXchPool.minCap = b->tilePartPool.minCap = 256;
    L(&d.lneg, 4, 1, false); L(&d.lpos, 4, 1, false); L(&d.lmeta, 4, 1, false); L(&d.lsim, 4, 1, false); L(&d.lflags, 4, 1, false);
    L(&d.q2x, R, 1, false); L(&d.q2y, R, 1, false); L(&d.q2z, R, 1, false);
    L(&d.a1x, R, 1, false); L(&d.a1y, R, 1, false); L(&d.a1z, R, 1, false);
    L(&d.a2x, R, 1, false); L(&d.a2y, R, 1, false); L(&d.a2z, R, 1, false);
    L(&d.strain, 4, 1, false); L(&d.maxStrain, 4, 1, false); L(&d.strainOffset, 4, 1, false); L(&d.lnew, 4, 1, false);
}

// room for bytes in the staging buffer: waits for the main stream (every copy from the
// buffer has run) when the buffer is full, and grows it when it is too small
vx3h_status stageAlloc(vx3h_batch* b, size_t bytes, char** out) {
    const size_t r = (bytes + 255) & ~(size_t)255;
    if (b->stageUsed + r > b->stageCap) {
        HIPCHK(hipStreamSynchronize(b->stream));
        b->stageUsed = 0;
        if (r > b->stageCap) {
            if (b->stage) HIPCHK(hipHostFree(b->stage));
            b->stage = nullptr;
            b->stageCap = 0;
            const size_t c = std::max(r, std::max(2 * b->stageCap, (size_t)16 << 20));
            HIPCHK(hipHostMalloc((void**)&b->stage, c, hipHostMallocDefault));
            b->stageCap = c;
        }
    }
    *out = b->stage + b->stageUsed;
    b->stageUsed += r;
    return VX3H_OK;
}

// host-to-device copy on the main stream through the staging buffer (the source may be
// freed on return)
vx3h_status upload(vx3h_batch* b, void* dst, const void* src, size_t bytes) {
    if (!bytes) return VX3H_OK;
    char* p;
    if (vx3h_status st = stageAlloc(b, bytes, &p)) return st;
    memcpy(p, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, b->stream));
    return VX3H_OK;
}

vx3h_status fillBytes(vx3h_batch* b, void* dst, int value, size_t bytes) {
    if (bytes) HIPCHK(hipMemsetAsync(dst, value, bytes, b->stream));
    return VX3H_OK;
}

vx3h_status growArrays(vx3h_batch* b, std::vector<ArrSpec>& arrs, int& cap, int need, int used, bool isVoxel) {
    if (need <= cap) return VX3H_OK;
    int ncap = std::max(need, std::max(2 * cap, 1024));
    HIPCHK(hipStreamSynchronize(b->stream));
    // The arrays are replaced one at a time (the old one freed after its copy), so the
    // device needs the growth of all of them plus the largest new one.  Checked first: an
    // allocation failing half way would leave arrays of two capacities (and two strides).
    {
        size_t grow = 0, largest = 0, freeB = 0, totalB = 0;
        for (const ArrSpec& a : arrs) {
            grow += (size_t)(ncap - cap) * a.per * a.elem;
            largest = std::max(largest, (size_t)ncap * a.per * a.elem);
        }
        HIPCHK(hipMemGetInfo(&freeB, &totalB));
        if (grow + largest > freeB) FAILV(VX3H_ENOMEM, "device memory too small to grow the batch arrays");
    }
    for (ArrSpec& a : arrs) {
        void* np = nullptr;
        size_t bytes = (size_t)ncap * a.per * a.elem;
        HIPCHK(hipMalloc(&np, bytes));
        HIPCHK(hipMemsetAsync(np, 0, bytes, b->stream));
        if (*a.slot && used > 0) {
            if (a.strided)
                HIPCHK(hipMemcpy2DAsync(np, (size_t)ncap * a.elem, *a.slot, (size_t)cap * a.elem, (size_t)used * a.elem, a.per,
                                        hipMemcpyDeviceToDevice, b->stream));
            else
                HIPCHK(hipMemcpyAsync(np, *a.slot, (size_t)used * a.per * a.elem, hipMemcpyDeviceToDevice, b->stream));
        }
        HIPCHK(hipStreamSynchronize(b->stream));
        if (*a.slot) HIPCHK(hipFree(*a.slot));
        *a.slot = np;
    }
    // unused elements belong to no simulation
    int* owner = isVoxel ? b->db.vsim : b->db.lsim;
    if (ncap > used) HIPCHK(hipMemsetAsync(owner + used, 0xFF, (size_t)(ncap - used) * 4, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    cap = ncap;
    if (isVoxel) b->db.Fs = (size_t)ncap;
    b->layout++;
    return VX3H_OK;
}

vx3h_status poolAlloc(vx3h_batch* b, DevPool& p, int n, int& start) {
    bool grow = false;
    start = p.ra.alloc(n, p.cap, grow);
    if (!grow) return VX3H_OK;
    const int ncap = std::max(p.ra.hi, std::max(2 * p.cap, p.minCap));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (auto& a : p.arrs) {
        void* np = nullptr;
        HIPCHK(hipMalloc(&np, (size_t)ncap * a.second));
        HIPCHK(hipMemsetAsync(np, 0xFF, (size_t)ncap * a.second, b->stream));   // hashSim: no owner
        if (*a.first && p.cap)
            HIPCHK(hipMemcpyAsync(np, *a.first, (size_t)p.cap * a.second, hipMemcpyDeviceToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        if (*a.first) HIPCHK(hipFree(*a.first));
        *a.first = np;
    }
    p.cap = ncap;
    b->layout++;
    return VX3H_OK;
}

template <typename E> vx3h_status uploadTable(vx3h_batch* b, void*& dptr, size_t& cap, const std::vector<E>& v) {
    if (v.empty()) return VX3H_OK;
    if (v.size() > cap) {
        size_t nc = std::max(v.size(), 2 * cap + 16);
        void* np = nullptr;
        HIPCHK(hipStreamSynchronize(b->stream));
        HIPCHK(hipMalloc(&np, nc * sizeof(E)));
        if (dptr) HIPCHK(hipFree(dptr));
        dptr = np;
        cap = nc;
        b->layout++;
    }
    return upload(b, dptr, v.data(), v.size() * sizeof(E));
}

template <typename T> void putReal(std::vector<char>& out, const std::vector<double>& src) {
    out.resize(src.size() * sizeof(T));
    T* p = (T*)out.data();
    for (size_t i = 0; i < src.size(); i++) p[i] = (T)src[i];
}

// reals in the state precision, converted straight into the staging buffer
vx3h_status copyReal(vx3h_batch* b, void* dst, const std::vector<double>& src) {
    const size_t bytes = src.size() * b->realSize;
    if (!bytes) return VX3H_OK;
    char* p;
    if (vx3h_status st = stageAlloc(b, bytes, &p)) return st;
    if (b->realSize == 4)
        for (size_t i = 0; i < src.size(); i++) ((float*)p)[i] = (float)src[i];
    else
        memcpy(p, src.data(), bytes);
    HIPCHK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, b->stream));
    return VX3H_OK;
}

vx3h_status readReal(vx3h_batch* b, const void* src, size_t n, std::vector<double>& out) {
    out.resize(n);
    if (b->realSize == 4) {
        std::vector<float> t(n);
        HIPCHK(hipMemcpy(t.data(), src, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) out[i] = t[i];
    } else {
        HIPCHK(hipMemcpy(out.data(), src, n * 8, hipMemcpyDeviceToHost));
    }
    return VX3H_OK;
}

char* at(void* base, size_t elem, size_t idx) { return (char*)base + elem * idx; }

constexpr NearBitTable kHostNearBits;

// Uploads the nearby sets of voxels [v0, v0 + nv) as lattice-offset bit masks (SC_NEARMASK):
// voxel v's set is list[start[v], start[v] + count[v]) of sim-local indices, ijk the
// sim's lattice indices (3 per voxel).
vx3h_status uploadNearMasks(vx3h_batch* b, int v0, int nv, const int* ijk, const std::vector<int>& start,
                            const std::vector<int>& count, const std::vector<int>& list) {
    std::vector<unsigned> m((size_t)VX3_NEAR_WORDS * nv, 0u);
    for (int v = 0; v < nv; v++)
        for (int i = 0; i < count[v]; i++) {
            const int u = list[start[v] + i];
            const int dx = ijk[3 * u] - ijk[3 * v], dy = ijk[3 * u + 1] - ijk[3 * v + 1], dz = ijk[3 * u + 2] - ijk[3 * v + 2];
            if (dx < -5 || dx > 5 || dy < -5 || dy > 5 || dz < -5 || dz > 5) FAILV(VX3H_EINVAL, "nearby voxel beyond the mask's offsets");
            const int bit = kHostNearBits.bit[(dx + 5) * 121 + (dy + 5) * 11 + (dz + 5)];
            if (bit < 0) FAILV(VX3H_EINVAL, "nearby voxel beyond the mask's offsets");
            m[(size_t)(bit >> 5) * nv + v] |= 1u << (bit & 31);
        }
    if (nv > 0)
        for (int w = 0; w < VX3_NEAR_WORDS; w++)
            if (vx3h_status st = upload(b, b->db.nearMask + (size_t)w * b->db.Vcap + v0, m.data() + (size_t)w * nv, (size_t)nv * 4))
                return st;
    return VX3H_OK;
}

StepOpts stepOpts(const vx3h_batch* b);
vx3h_status placeLoads(vx3h_batch* b, bool scratch);

vx3h_status launchSteps(vx3h_batch* b, int n) {
    const StepOpts o = stepOpts(b);
    for (int i = 0; i < n; i++) {
        if (b->realSize == 4) launchStep(b->db, b->stream, o);
        else launchStep(*reinterpret_cast<DevBatch<double>*>(&b->db), b->stream, o);
    }
    HIPCHK(hipGetLastError());
    return VX3H_OK;
}

vx3h_status ensureGraph(vx3h_batch* b, int steps) {
    if (b->graphExec && b->graphLayout == b->layout && b->graphSteps == steps) return VX3H_OK;
    if (b->graphExec) { HIPCHK(hipGraphExecDestroy(b->graphExec)); b->graphExec = nullptr; }
    if (b->graph) { HIPCHK(hipGraphDestroy(b->graph)); b->graph = nullptr; }
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    vx3h_status st = launchSteps(b, steps);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st != VX3H_OK) return st;
    if (e != hipSuccess) FAILV(VX3H_EDEVICE, std::string("graph capture failed: ") + hipGetErrorString(e));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->graphExec, g, nullptr, nullptr, 0));
    b->graphLayout = b->layout;
    b->graphSteps = steps;
    return VX3H_OK;
}

vx3h_status syncResLists(vx3h_batch* b) {
    if (!b->resDirty) return VX3H_OK;
    for (size_t i = 0; i < b->resLists.size(); i++) {
        const auto& L = b->resLists[i];
        if (L.empty()) continue;
        if (L.size() > b->capRes[i]) {
            if (b->dResLists[i]) HIPCHK(hipFree(b->dResLists[i]));
            b->capRes[i] = std::max(L.size(), 2 * b->capRes[i] + 64);
            HIPCHK(hipMalloc(&b->dResLists[i], b->capRes[i] * sizeof(int)));
        }
        HIPCHK(hipMemcpy(b->dResLists[i], L.data(), L.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    b->resDirty = false;
    return VX3H_OK;
}

// grows a device array to at least n elements (contents not kept)
template <typename T> vx3h_status growDev(T** p, size_t* cap, size_t n) {
    if (n <= *cap) return VX3H_OK;
    if (*p) HIPCHK(hipFree(*p));
    *p = nullptr;
    const size_t c = std::max(n, *cap + *cap / 2);
    HIPCHK(hipMalloc((void**)p, c * sizeof(T)));
    *cap = c;
    return VX3H_OK;
}

// One-tile tables of a one-workgroup simulation (its voxels and links in order, no halo,
// no exchange records), uploaded once.
vx3h_status ensureOneTile(vx3h_batch* b, int s) {
    SlotHost& sh = b->slots[s];
    if (sh.oneTile) return VX3H_OK;
    TileDesc d{};
    d.tile = 0; d.nTiles = 1; d.vOff = 0; d.nOwn = sh.nv; d.nHalo = 0; d.lOff = 0; d.nLinks = sh.nl;
    std::vector<int> vox(sh.nv), xidx(sh.nv, -1), link(sh.nl), ends(sh.nl);
    for (int i = 0; i < sh.nv; i++) vox[i] = i;
    for (int l = 0; l < sh.nl; l++) {
        link[l] = l | (int)0x80000000u;
        ends[l] = sh.lneg[l] | (sh.lpos[l] << 16);
    }
    sh.tvN = std::max(1, sh.nv);
    sh.tlN = std::max(1, sh.nl);
    sh.txN = 1;
    vx3h_status st;
    if ((st = poolAlloc(b, b->tileSlotPool, sh.tvN, sh.tv0)) || (st = poolAlloc(b, b->tileLinkPool, sh.tlN, sh.tl0)) ||
        (st = poolAlloc(b, b->tileXchPool, sh.txN, sh.tx0)) || (st = poolAlloc(b, b->tilePartPool, 1, sh.tp0)))
        return st;
    if ((st = upload(b, b->dTileVox + sh.tv0, vox.data(), vox.size() * 4)) ||
        (st = upload(b, b->dTileXidx + sh.tv0, xidx.data(), xidx.size() * 4)) ||
        (st = upload(b, b->dTileLink + sh.tl0, link.data(), link.size() * 4)) ||
        (st = upload(b, b->dTileEnds + sh.tl0, ends.data(), ends.size() * 4)))
        return st;
    sh.tplan.tiles.assign(1, d);
    sh.tplan.nPub = 0;
    sh.oneTile = true;
    return VX3H_OK;
}

// The tiled launch groups, after a simulation joined or left: whole simulations packed
// into groups of at most tileCap tiles (first fit, largest first: a launch lasts as long
// per step whether full or not), and their tile descriptors pointing at each
// simulation's ranges in the tile pools (filled at admission).  A group's sync area
// holds one divergence word per simulation, then one arrival flag per workgroup.
vx3h_status syncTiles(vx3h_batch* b) {
    if (!b->tileDirty) return VX3H_OK;
    std::vector<int> order = b->tiledSims;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->slots[x].tiles > b->slots[y].tiles; });
    // with tiled simulations running, the one-workgroup simulations without collisions
    // (lists [0, shapes)) fill the launches' free slots as one-tile entries, when every one
    // of them fits the tiled kernel's shape
    b->absorbed = !order.empty();
    const ResidentShapeInfo& ts = b->shapes[b->tiledShape];
    for (size_t i = 0; b->absorbed && i < b->shapes.size(); i++)
        for (int s : b->resLists[i])
            if (b->slots[s].nv > ts.block * ts.vpt || b->slots[s].axisLinks > ts.block * ts.lpa) b->absorbed = false;
    if (b->absorbed)
        for (size_t i = 0; i < b->shapes.size(); i++)
            for (int s : b->resLists[i]) {
                vx3h_status st = ensureOneTile(b, s);
                if (st) return st;
                order.push_back(s);
            }
    auto tilesOf = [&](int s) { return b->slots[s].tiles ? b->slots[s].tiles : 1; };
    std::vector<std::vector<int>> bins;
    std::vector<int> fill;
    for (int s : order) {
        size_t i = 0;
        while (i < bins.size() && fill[i] + tilesOf(s) > b->tileCap) i++;
        if (i == bins.size()) { bins.emplace_back(); fill.push_back(0); }
        bins[i].push_back(s);
        fill[i] += tilesOf(s);
    }
    std::vector<TileDesc> T;
    b->tileGroups.clear();
    int syncTot = 0;
    for (const auto& bin : bins) {
        vx3h_batch::TileGroup g{(int)T.size(), 0, syncTot, 0};
        for (int s : bin) {
            const SlotHost& sh = b->slots[s];
            for (TileDesc d : sh.tplan.tiles) {
                d.sim = s;
                d.vOff += sh.tv0;
                d.lOff += sh.tl0;
                d.xOff = sh.tx0 * 2 * VX3_XCH_STRIDE;
                d.xPar = VX3_XCH_STRIDE * sh.tplan.nPub;
                d.cOff = g.nSync;
                d.pOff = sh.tp0 * 8;
                T.push_back(d);
            }
            g.count += tilesOf(s);
            g.nSync++;
            syncTot++;
        }
        syncTot += g.count;   // the arrival flags, one per workgroup, after the divergence words
        b->tileGroups.push_back(g);
    }
    vx3h_status st;
    if ((st = growDev(&b->dTiles, &b->capTiles, T.size())) || (st = growDev(&b->dSync, &b->capSync, (size_t)syncTot))) return st;
    if (!T.empty()) HIPCHK(hipMemcpy(b->dTiles, T.data(), T.size() * sizeof(TileDesc), hipMemcpyHostToDevice));
    b->tileDirty = false;
    return VX3H_OK;
}

vx3h_status takeEvent(vx3h_batch* b, hipEvent_t& e) {
    if (b->evPool.empty()) { HIPCHK(hipEventCreate(&e)); }
    else { e = b->evPool.back(); b->evPool.pop_back(); }
    return VX3H_OK;
}

// Launches every resident shape's kernel for nSteps steps, each on its own stream
// after the main stream's work so far; joinResident makes the main stream wait for
// them.  Between the two the caller enqueues the stream-path steps on the main stream.
vx3h_status launchResidentAll(vx3h_batch* b, int nSteps, hipEvent_t* ev) {
    vx3h_status st = syncResLists(b);
    if (st || (st = syncTiles(b))) return st;
    bool forked = false;
    for (size_t i = 0; i < b->resLists.size(); i++) {
        const int n = (int)b->resLists[i].size();
        if (!n || (b->absorbed && i < b->shapes.size())) continue;   // absorbed: they run in the tiled launches
        if (!forked) {
            if (!b->forkEv) HIPCHK(hipEventCreateWithFlags(&b->forkEv, hipEventDisableTiming));
            HIPCHK(hipEventRecord(b->forkEv, b->stream));
            forked = true;
        }
        if (!b->resStreams[i]) {
            HIPCHK(hipStreamCreateWithFlags(&b->resStreams[i], hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&b->joinEv[i], hipEventDisableTiming));
        }
        HIPCHK(hipStreamWaitEvent(b->resStreams[i], b->forkEv, 0));
        b->joinPending[i] = true;
        hipEvent_t pair[2];
        hipEvent_t* evp = ev ? ev + 2 * i : nullptr;
        if (!ev && b->timing) {
            if ((st = takeEvent(b, pair[0])) || (st = takeEvent(b, pair[1]))) return st;
            b->evUsed.push_back({pair[0], pair[1]});
            b->timedSteps += nSteps;
            evp = pair;
        }
        hipStream_t rs = b->resStreams[i];
        const int ns = (int)b->shapes.size();
        const int shape = (int)i % ns;
        const bool col = (int)i >= ns;
        hipError_t e = b->realSize == 4
            ? launchResident(b->db, rs, shape, col, b->dResLists[i], n, nSteps, evp)
            : launchResident(*reinterpret_cast<DevBatch<double>*>(&b->db), rs, shape, col, b->dResLists[i], n, nSteps, evp);
        if (e != hipSuccess) FAILV(VX3H_EDEVICE, std::string("resident launch failed: ") + hipGetErrorString(e));
    }
    // tiled simulations: their launch groups one after another on the tile stream
    if (!b->tileGroups.empty()) {
        if (!forked) {
            if (!b->forkEv) HIPCHK(hipEventCreateWithFlags(&b->forkEv, hipEventDisableTiming));
            HIPCHK(hipEventRecord(b->forkEv, b->stream));
        }
        if (!b->tileStream) {
            HIPCHK(hipStreamCreateWithFlags(&b->tileStream, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&b->tileJoin, hipEventDisableTiming));
        }
        HIPCHK(hipStreamWaitEvent(b->tileStream, b->forkEv, 0));
        b->tilePending = true;
        for (size_t gi = 0; gi < b->tileGroups.size(); gi++) {
            const vx3h_batch::TileGroup& g = b->tileGroups[gi];
            TileArgs ta{};
            ta.tiles = b->dTiles + g.first;
            ta.vox = b->dTileVox;
            ta.link = b->dTileLink;
            ta.ends = b->dTileEnds;
            ta.xidx = b->dTileXidx;
            ta.xch = b->dXch;
            ta.sync = b->dSync + g.sync0;
            ta.flag0 = g.nSync;
            ta.part = b->dPart;
            ta.spin = b->spinTicks;
            hipEvent_t pair[2];
            hipEvent_t* evp = ev ? ev + 2 * (b->resLists.size() + gi) : nullptr;
            if (!ev && b->timing) {
                if ((st = takeEvent(b, pair[0])) || (st = takeEvent(b, pair[1]))) return st;
                b->evUsed.push_back({pair[0], pair[1]});
                b->timedSteps += nSteps;
                evp = pair;
            }
            const hipError_t e = b->realSize == 4
                ? launchTiled(b->db, b->tileStream, b->tiledShape, ta, g.count, g.nSync + g.count, nSteps, evp)
                : launchTiled(*reinterpret_cast<DevBatch<double>*>(&b->db), b->tileStream, b->tiledShape, ta, g.count,
                              g.nSync + g.count, nSteps, evp);
            if (e != hipSuccess) FAILV(VX3H_EDEVICE, std::string("tiled resident launch failed: ") + hipGetErrorString(e));
        }
    }
    return VX3H_OK;
}

vx3h_status joinResident(vx3h_batch* b) {
    for (size_t i = 0; i < b->resStreams.size(); i++) {
        if (!b->joinPending[i]) continue;
        HIPCHK(hipEventRecord(b->joinEv[i], b->resStreams[i]));
        HIPCHK(hipStreamWaitEvent(b->stream, b->joinEv[i], 0));
        b->joinPending[i] = false;
    }
    if (b->tilePending) {
        HIPCHK(hipEventRecord(b->tileJoin, b->tileStream));
        HIPCHK(hipStreamWaitEvent(b->stream, b->tileJoin, 0));
        b->tilePending = false;
    }
    return VX3H_OK;
}

}  // namespace

extern "C" vx3h_status vx3h_batch_create(int device, const vx3h_batch_cfg* cfg, vx3h_batch** out) {
    if (!out) FAILV(VX3H_EINVAL, "null output handle");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) FAILV(VX3H_EDEVICE, "no HIP device available");
    if (device < 0 || device >= n) FAILV(VX3H_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    vx3h_batch* b = new vx3h_batch();
    b->device = device;
    if (cfg) b->cfg = *cfg;
    b->realSize = b->cfg.precision == VX3H_PREC_F64 ? 8 : 4;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete b; FAILV(VX3H_EDEVICE, "stream creation failed"); }
    memset(&b->db, 0, sizeof(b->db));
    b->db.solo = -1;
    if (hipMalloc(&b->db.colAny, 256) != hipSuccess || hipMemset(b->db.colAny, 0, 256) != hipSuccess) {
        hipStreamDestroy(b->stream);
        delete b;
        FAILV(VX3H_EDEVICE, "device allocation failed");
    }
    initSpecs(b);
    {
        ResidentShapeInfo tmp[16];
        int ns = residentShapes(b->realSize == 8, tmp, 16);
        b->shapes.assign(tmp, tmp + ns);
        // list i < ns: shape i without collisions; ns + i: shape i with collisions
        b->resLists.assign(2 * ns, {});
        b->dResLists.assign(2 * ns, nullptr);
        b->capRes.assign(2 * ns, 0);
        b->resStreams.assign(2 * ns, nullptr);
        b->joinEv.assign(2 * ns, nullptr);
        b->joinPending.assign(2 * ns, false);
        // default preference: smallest voxel capacity first (more workgroups per CU), table order on ties
        for (int i = 0; i < ns; i++) b->shapePref.push_back(i);
        std::stable_sort(b->shapePref.begin(), b->shapePref.end(), [&](int x, int y) {
            return b->shapes[x].block * b->shapes[x].vpt < b->shapes[y].block * b->shapes[y].vpt;
        });
        // a configured block size moves the shapes of that block size to the front
        if (b->cfg.block_size > 0)
            std::stable_sort(b->shapePref.begin(), b->shapePref.end(), [&](int x, int y) {
                return (b->shapes[x].block == b->cfg.block_size) > (b->shapes[y].block == b->cfg.block_size);
            });
        // tiles take the preferred shape of the largest voxel capacity; a launch holds as many
        // tiles as the device keeps resident at once (a cooperative launch)
        int vmaxBest = 0;
        for (int si : b->shapePref) vmaxBest = std::max(vmaxBest, b->shapes[si].block * b->shapes[si].vpt);
        for (int si : b->shapePref)
            if (b->shapes[si].block * b->shapes[si].vpt == vmaxBest) { b->tiledShape = si; break; }
        int coop = 0, cus = 0, perCU = 0, khz = 0;
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device);
        const hipError_t oe = b->realSize == 4 ? tiledOccupancy(b->db, b->tiledShape, &perCU)
                                               : tiledOccupancy(*reinterpret_cast<DevBatch<double>*>(&b->db), b->tiledShape, &perCU);
        b->tileCap = (coop && oe == hipSuccess && khz > 0) ? perCU * cus : 0;
        b->spinTicks = (long long)khz * 1000 * 4;   // 4 s
    }
    size_t freeB = 0, totB = 0;
    if (hipMemGetInfo(&freeB, &totB) != hipSuccess) freeB = 0;
    b->budget = b->cfg.max_device_bytes ? b->cfg.max_device_bytes : (size_t)(0.9 * (double)freeB);
    *out = b;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_batch_destroy(vx3h_batch* b) {
    if (!b) return VX3H_OK;
    hipSetDevice(b->device);
    hipStreamSynchronize(b->stream);
    for (size_t i = 0; i < b->resStreams.size(); i++) {
        if (b->resStreams[i]) { hipStreamSynchronize(b->resStreams[i]); hipStreamDestroy(b->resStreams[i]); }
        if (b->joinEv[i]) hipEventDestroy(b->joinEv[i]);
    }
    if (b->tileStream) { hipStreamSynchronize(b->tileStream); hipStreamDestroy(b->tileStream); }
    if (b->tileJoin) hipEventDestroy(b->tileJoin);
    for (void* p : {(void*)b->dTiles, (void*)b->dSync}) if (p) hipFree(p);
    if (b->forkEv) hipEventDestroy(b->forkEv);
    if (b->graphExec) hipGraphExecDestroy(b->graphExec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (auto* arrs : {&b->vArrs, &b->lArrs})
        for (ArrSpec& a : *arrs) if (*a.slot) hipFree(*a.slot);
    for (void* p : {b->dVmat, b->dLmat, b->dExt, b->dTok, b->dSimMats, (void*)b->db.sims, (void*)b->db.hist}) if (p) hipFree(p);
    for (int* p : b->dResLists) if (p) hipFree(p);
    if (b->db.colAny) hipFree(b->db.colAny);
    for (DevPool* pl : {&b->nbPool, &b->hashPool, &b->nearPool, &b->cellPool, &b->tgtPool, &b->attPool, &b->pairPool,
                        &b->tileSlotPool, &b->tileLinkPool, &b->tileXchPool, &b->tilePartPool})
        for (auto& a : pl->arrs) if (*a.first) hipFree(*a.first);
    for (auto& pr : b->evUsed) { b->evPool.push_back(pr.first); b->evPool.push_back(pr.second); }
    for (hipEvent_t e : b->evPool) hipEventDestroy(e);
    if (b->stage) hipHostFree(b->stage);
    hipStreamDestroy(b->stream);
    delete b;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_sim_footprint(const vx3h_sim_desc* d, int precision, size_t* bytes) {
    if (!d || !bytes) FAILV(VX3H_EINVAL, "null argument");
    HostModel H;
    if (!buildModel(d, H, g_err)) return VX3H_EINVAL;
    size_t rs = precision == VX3H_PREC_F64 ? 8 : 4;
    const int NLcap = d->enable_attach && H.nv > 0 ? std::max(H.nl, 3 * H.nv) : H.nl;
    *bytes = (size_t)H.nv * voxelBytes(rs) + (size_t)NLcap * linkBytes(rs) + sizeof(SimCtl) + H.exts.size() * sizeof(ExtDev);
    return VX3H_OK;
}

namespace vx3 {
struct PreparedSim {
    HostModel H;
    long long nv = 0, nlCap = 0;   // voxels and link slots (room for attached links included)
    bool planned = false;   // the plan below is the tiled shape's (or none fits one launch)
    int vmax = 0, lax = 0, cap = 0;
    TilePlan plan;
    bool planOk = false;
};

// the tile plan a simulation of this model gets if it is too large for one workgroup
// (planned for any model over the tiled shape's one-workgroup capacity without collisions;
// admission decides whether to use it)
vx3h_status prepareSim(const vx3h_batch* b, const vx3h_sim_desc* desc, PreparedSim** out) {
    *out = nullptr;
    if (!b || !desc) FAILV(VX3H_EINVAL, "null argument");
    PreparedSim* p = new PreparedSim();
    if (!buildModel(desc, p->H, g_err)) { delete p; return VX3H_EINVAL; }
    const HostModel& H = p->H;
    p->nv = H.nv;
    p->nlCap = desc->enable_attach && H.nv > 0 ? std::max(H.nl, 3 * H.nv) : H.nl;
    if (b->tiledShape >= 0 && b->tileCap > 1 && !H.collisions && H.nv > 0) {
        const ResidentShapeInfo& r = b->shapes[b->tiledShape];
        int perAxis[3] = {0, 0, 0};
        for (int i = 0; i < H.nl; i++) perAxis[H.laxis[i]]++;
        const int lax = r.block * r.lpa, vmax = r.block * r.vpt;
        if (H.nv > vmax || perAxis[0] > lax || perAxis[1] > lax || perAxis[2] > lax) {
            p->planned = true;
            p->vmax = vmax; p->lax = lax; p->cap = b->tileCap;
            p->planOk = planTiles(H.nv, H.ijk.data(), H.nl, H.lneg.data(), H.lpos.data(), H.laxis.data(), vmax, lax,
                                  b->tileCap, &p->plan);
        }
    }
    *out = p;
    return VX3H_OK;
}

void freePrepared(PreparedSim* p) { delete p; }

void preparedSize(const PreparedSim* p, long long* nv, long long* nl) {
    *nv = p ? p->nv : 0;
    *nl = p ? p->nlCap : 0;
}

vx3h_status reserveBatch(vx3h_batch* b, long long nv, long long nl) {
    if (!b) FAILV(VX3H_EINVAL, "null batch");
    HIPCHK(hipSetDevice(b->device));
    vx3h_status st;
    const long long vneed = b->va.hi + nv, lneed = b->la.hi + nl;
    if (vneed > INT32_MAX || lneed > INT32_MAX) FAILV(VX3H_ENOMEM, "batch beyond 2^31 voxels or links");
    if (vneed > b->db.Vcap && (st = growArrays(b, b->vArrs, b->db.Vcap, (int)vneed, b->va.hi, true))) return st;
    if (lneed > b->db.Lcap && (st = growArrays(b, b->lArrs, b->db.Lcap, (int)lneed, b->la.hi, false))) return st;
    return VX3H_OK;
}
}  // namespace vx3

extern "C" vx3h_status vx3h_sim_admit(vx3h_batch* b, const vx3h_sim_desc* desc, int* slot) {
    return vx3::admitPrepared(b, desc, nullptr, slot);
}

namespace vx3 {
// the admission's copies and fills, on the main stream
#define UPLOAD(...) do { if (vx3h_status st_ = upload(b, __VA_ARGS__)) return st_; } while (0)
#define FILL(...) do { if (vx3h_status st_ = fillBytes(b, __VA_ARGS__)) return st_; } while (0)

vx3h_status admitPrepared(vx3h_batch* b, const vx3h_sim_desc* desc, PreparedSim* pre, int* slot) {
    std::unique_ptr<PreparedSim> own(pre);
    if (!b || !desc || !slot) FAILV(VX3H_EINVAL, "null argument");
    if (desc->stop_condition_type < VX3H_SC_NONE || desc->stop_condition_type > VX3H_SC_MIN_MAXMOVE)
        FAILV(VX3H_EINVAL, "unknown stop condition type");
    HIPCHK(hipSetDevice(b->device));
    HostModel Hbuilt;
    if (!pre && !buildModel(desc, Hbuilt, g_err)) return VX3H_EINVAL;
    HostModel& H = pre ? pre->H : Hbuilt;
    if (H.stopTok.size() > 64) FAILV(VX3H_EUNSUPPORTED, "stop condition longer than 64 tokens");
    for (int k = 0; k < 3; k++)
        if (H.ffTok[k].size() > 64) FAILV(VX3H_EUNSUPPORTED, "force-field program longer than 64 tokens");
    for (int k = 0; k < 5; k++)
        if (H.atTok[k].size() > 64) FAILV(VX3H_EUNSUPPORTED, "attach-condition program longer than 64 tokens");
    const bool attach = desc->enable_attach != 0 && H.nv > 0;
    // attach needs room for new links: every voxel has six faces, so at most 3 links per voxel
    const int NLcap = attach ? std::max(H.nl, 3 * H.nv) : H.nl;
    const size_t rs = b->realSize;
    const size_t need = (size_t)H.nv * voxelBytes(rs) + (size_t)NLcap * linkBytes(rs) + sizeof(SimCtl) + H.exts.size() * sizeof(ExtDev);
    if (b->bytesUsed + need > b->budget) FAILV(VX3H_ENOMEM, "simulation does not fit the batch's device memory budget");
    HIPCHK(hipStreamSynchronize(b->stream));
    b->stageUsed = 0;   // the stream has run every staged copy
    // slot
    int s = -1;
    for (int i = 0; i < (int)b->slots.size(); i++) if (!b->slots[i].used) { s = i; break; }
    if (s < 0) { s = (int)b->slots.size(); b->slots.emplace_back(); }
    if ((int)b->slots.size() > b->simCap) {
        int nc = std::max((int)b->slots.size(), 2 * b->simCap + 8);
        SimCtl* np = nullptr;
        HIPCHK(hipMalloc(&np, nc * sizeof(SimCtl)));
        FILL(np, 0, nc * sizeof(SimCtl));
        double* nh = nullptr;
        const size_t hs = 3 * VX3_HIST * sizeof(double);
        HIPCHK(hipMalloc(&nh, nc * hs));
        if (b->db.sims) {
            HIPCHK(hipMemcpyAsync(np, b->db.sims, b->simCap * sizeof(SimCtl), hipMemcpyDeviceToDevice, b->stream));
            HIPCHK(hipMemcpyAsync(nh, b->db.hist, b->simCap * hs, hipMemcpyDeviceToDevice, b->stream));
            HIPCHK(hipStreamSynchronize(b->stream));
            HIPCHK(hipFree(b->db.sims));
            HIPCHK(hipFree(b->db.hist));
        }
        b->db.sims = np;
        b->db.hist = nh;
        b->simCap = nc;
        b->layout++;
    }
    // ranges
    bool gv = false, gl = false;
    int v0 = H.nv ? b->va.alloc(H.nv, b->db.Vcap, gv) : 0;
    int l0 = NLcap ? b->la.alloc(NLcap, b->db.Lcap, gl) : 0;
    if (gv) { vx3h_status st = growArrays(b, b->vArrs, b->db.Vcap, b->va.hi, v0, true); if (st) return st; }
    if (gl) { vx3h_status st = growArrays(b, b->lArrs, b->db.Lcap, b->la.hi, l0, false); if (st) return st; }
    // tables
    std::vector<int> vmIdx(H.vmats.size()), lmIdx(H.lmats.size());
    size_t nvm = b->vmats.items.size(), nlm = b->lmats.items.size();
    for (size_t i = 0; i < H.lmats.size(); i++) lmIdx[i] = b->lmats.add(H.lmats[i]);
    for (size_t i = 0; i < H.vmats.size(); i++) {
        if (H.vmSelf[i] >= 0) H.vmats[i].selfLink = lmIdx[H.vmSelf[i]];
        vmIdx[i] = b->vmats.add(H.vmats[i]);
    }
    if (b->vmats.items.size() != nvm || !b->dVmat) { vx3h_status st = uploadTable(b, b->dVmat, b->capVmat, b->vmats.items); if (st) return st; }
    if (b->lmats.items.size() != nlm || !b->dLmat) { vx3h_status st = uploadTable(b, b->dLmat, b->capLmat, b->lmats.items); if (st) return st; }
    int ext0 = (int)b->exts.size();
    if (!H.exts.empty()) {
        b->exts.insert(b->exts.end(), H.exts.begin(), H.exts.end());
        vx3h_status st = uploadTable(b, b->dExt, b->capExt, b->exts);
        if (st) return st;
    }
    int mat0 = (int)b->simMats.size();
    b->simMats.insert(b->simMats.end(), vmIdx.begin(), vmIdx.end());
    b->simMats.insert(b->simMats.end(), lmIdx.begin(), lmIdx.end());
    { vx3h_status st = uploadTable(b, b->dSimMats, b->capSimMats, b->simMats); if (st) return st; }
    b->db.simMats = (const int*)b->dSimMats;
    int tok0 = (int)b->tokens.size();
    for (const vx3h_token& t : H.stopTok) { TokenDev td; td.op = t.op; td.pad = 0; td.value = t.value; b->tokens.push_back(td); }
    int ffOff[3], atOff[5];
    for (int k = 0; k < 3; k++) {
        ffOff[k] = (int)b->tokens.size();
        for (const vx3h_token& t : H.ffTok[k]) { TokenDev td; td.op = t.op; td.pad = 0; td.value = t.value; b->tokens.push_back(td); }
    }
    for (int k = 0; k < 5; k++) {
        atOff[k] = (int)b->tokens.size();
        for (const vx3h_token& t : H.atTok[k]) { TokenDev td; td.op = t.op; td.pad = 0; td.value = t.value; b->tokens.push_back(td); }
    }

    if ((int)b->tokens.size() > tok0) { vx3h_status st = uploadTable(b, b->dTok, b->capTok, b->tokens); if (st) return st; }
    b->db.vmatTab = (const VoxMatDev*)b->dVmat;
    b->db.lmatTab = (const LinkMatDev*)b->dLmat;
    b->db.extTab = (const ExtDev*)b->dExt;
    b->db.tokens = (const TokenDev*)b->dTok;

    // voxel initial state (CVX_Voxel::reset + CVoxelyze::addVoxel, VX_Voxel.cpp:48-57, Voxelyze.cpp:439-461)
    DevBatch<float>& d = b->db;
    const int NV = H.nv, NL = H.nl;
    if (NV) {
        std::vector<double> px(NV), py(NV), pz(NV), zero(NV, 0.0), one(NV, 1.0);
        for (int i = 0; i < NV; i++) {
            px[i] = H.ijk[3 * i] * H.voxSize; py[i] = H.ijk[3 * i + 1] * H.voxSize; pz[i] = H.ijk[3 * i + 2] * H.voxSize;
        }
        vx3h_status st;
        if ((st = copyReal(b, at(d.px, rs, v0), px))) return st;
        if ((st = copyReal(b, at(d.py, rs, v0), py))) return st;
        if ((st = copyReal(b, at(d.pz, rs, v0), pz))) return st;
        for (void* p : {(void*)d.mx, (void*)d.my, (void*)d.mz, (void*)d.ox, (void*)d.oy, (void*)d.oz, (void*)d.ax, (void*)d.ay, (void*)d.az})
            if ((st = copyReal(b, at(p, rs, v0), zero))) return st;
        if (desc->init_velocity) {   // linear momentum = mass * velocity (the voxel material's float mass)
            void* mom[3] = {(void*)d.mx, (void*)d.my, (void*)d.mz};
            std::vector<double> lm(NV);
            for (int k = 0; k < 3; k++) {
                for (int i = 0; i < NV; i++) lm[i] = desc->init_velocity[3 * i + k] * (double)H.vmats[H.vmat[i]].mass;
                if ((st = copyReal(b, at(mom[k], rs, v0), lm))) return st;
            }
        }
        if ((st = copyReal(b, at(d.ow, rs, v0), one))) return st;
        std::vector<int> vm(NV), vs(NV, s), ve(NV);
        for (int i = 0; i < NV; i++) { vm[i] = vmIdx[H.vmat[i]]; ve[i] = H.vext[i] >= 0 ? ext0 + H.vext[i] : -1; }
        std::vector<int> vf(H.vflags);
        for (int i = 0; i < NV; i++)   // target materials: contacts counted as hits, close pairs
            if (desc->materials[desc->material[i]].is_target) vf[i] |= VF_TARGET;
        if (attach)   // VX3_Voxel::enableAttach starts true (VX3_Voxel.h:302)
            for (int i = 0; i < NV; i++) vf[i] |= VF_ATTACH;
        UPLOAD(d.vflags + v0, vf.data(), NV * 4);
        UPLOAD(d.tfac + v0, H.tfac0.data(), NV * 4);
        UPLOAD(d.phase + v0, H.phase.data(), NV * 8);
        UPLOAD(d.vmat + v0, vm.data(), NV * 4);
        UPLOAD(d.vext + v0, ve.data(), NV * 4);
        UPLOAD(d.vijk + 3 * (size_t)v0, H.ijk.data(), NV * 12);
        FILL(d.psx + v0, 0, NV * 4);
        FILL(d.psy + v0, 0, NV * 4);
        FILL(d.psz + v0, 0, NV * 4);
        for (int k = 0; k < 6; k++) FILL(d.es + (size_t)k * d.Vcap + v0, 0, NV * 4);
        UPLOAD(d.vsim + v0, vs.data(), NV * 4);
        if (!H.cilia.empty()) {   // per voxel: base, shift, the material's Cilia scale and switch-on time
            std::vector<double> col(NV);
            for (int k = 0; k < 8; k++) {
                for (int i = 0; i < NV; i++)
                    col[i] = k < 6 ? H.cilia[6 * (size_t)i + k] : k == 6 ? H.vmCilia[H.vmat[i]] : H.vmCiliaOn[H.vmat[i]];
                UPLOAD(d.cil + (size_t)k * d.Vcap + v0, col.data(), NV * 8);
            }
        }
    }
    if (NL) {
        std::vector<int> ln(NL), lp(NL), lm(NL), ls(NL, s), lf(NL, LF_SMALL_ANGLE);
        for (int i = 0; i < NL; i++) {
            ln[i] = v0 + H.lneg[i]; lp[i] = v0 + H.lpos[i];
            lm[i] = H.laxis[i] | (lmIdx[H.lmat[i]] << 2);
        }
        UPLOAD(d.lneg + l0, ln.data(), NL * 4);
        UPLOAD(d.lpos + l0, lp.data(), NL * 4);
        UPLOAD(d.lmeta + l0, lm.data(), NL * 4);
        UPLOAD(d.lflags + l0, lf.data(), NL * 4);
        std::vector<double> zero(NL, 0.0);
        vx3h_status st;
        for (void* p : {(void*)d.q2x, (void*)d.q2y, (void*)d.q2z, (void*)d.a1x, (void*)d.a1y, (void*)d.a1z, (void*)d.a2x, (void*)d.a2y, (void*)d.a2z})
            if ((st = copyReal(b, at(p, rs, l0), zero))) return st;
        FILL(d.strain + l0, 0, NL * 4);
        FILL(d.maxStrain + l0, 0, NL * 4);
        FILL(d.strainOffset + l0, 0, NL * 4);
        UPLOAD(d.lsim + l0, ls.data(), NL * 4);
    }
    if (NLcap) FILL(d.lnew + l0, 0, (size_t)NLcap * 4);
    if (NLcap > NL) {   // room for attached links: not owned until made
        FILL(d.lsim + l0 + NL, 0xFF, (size_t)(NLcap - NL) * 4);
        FILL(d.lflags + l0 + NL, 0, (size_t)(NLcap - NL) * 4);
    }
    // execution path: resident when the sim fits one of the on-chip kernel shapes
    int shape = -1;
    int perAxis[3] = {0, 0, 0};
    for (int i = 0; i < NL; i++) perAxis[H.laxis[i]]++;
    // the resident kernel evaluates the actuation once per distinct phase offset (bit pattern)
    std::vector<uint64_t> phases(NV);
    for (int i = 0; i < NV; i++) memcpy(&phases[i], &H.phase[i], 8);
    std::sort(phases.begin(), phases.end());
    const int nPhases = (int)(std::unique(phases.begin(), phases.end()) - phases.begin());
    // VoxCAD stop conditions and equilibrium mode are evaluated by the stream path's control pass
    const bool voxcad = desc->stop_condition_type != VX3H_SC_NONE || desc->equilibrium_mode != 0;
    const bool closePairs = desc->max_dist_close_pairs > 0;
    bool targets = false;   // target contacts (hit counts) are counted by the per-step kernels
    if (H.collisions)
        for (int i = 0; i < desc->n_materials; i++) targets |= desc->materials[i].is_target != 0;
    const bool resOk = (int)H.vmats.size() <= RES_MAX_VMATS && (int)H.lmats.size() <= RES_MAX_LMATS &&
                       nPhases <= RES_MAX_PHASES && !voxcad && !H.extras() && !desc->enable_detach &&
                       !desc->secondary_experiment && !desc->enable_signals && !attach && !closePairs && !targets;
    if (b->cfg.path != VX3H_PATH_STREAM && NV > 0 && resOk) {
        for (int si : b->shapePref) {
            const ResidentShapeInfo& r = b->shapes[si];
            const int lax = r.block * r.lpa;
            if (NV <= r.block * r.vpt && perAxis[0] <= lax && perAxis[1] <= lax && perAxis[2] <= lax) { shape = si; break; }
        }
    }
    // too large for one workgroup: tiles of the tiled resident kernel, when they fit one launch
    TilePlan tplan;
    int tiles = 0;
    if (b->cfg.path != VX3H_PATH_STREAM && NV > 0 && resOk && shape < 0 && !H.collisions && b->tiledShape >= 0 &&
        b->tileCap > 1) {
        const ResidentShapeInfo& r = b->shapes[b->tiledShape];
        bool ok;
        if (pre && pre->planned && pre->vmax == r.block * r.vpt && pre->lax == r.block * r.lpa && pre->cap == b->tileCap) {
            ok = pre->planOk;   // planned ahead for this shape
            tplan = std::move(pre->plan);
        } else
            ok = planTiles(NV, H.ijk.data(), NL, H.lneg.data(), H.lpos.data(), H.laxis.data(), r.block * r.vpt, r.block * r.lpa,
                           b->tileCap, &tplan);
        if (ok && tplan.tiles.size() > 1) tiles = (int)tplan.tiles.size();
    }
    if (b->cfg.path == VX3H_PATH_RESIDENT && NV > 0 && resOk && shape < 0 && !tiles)
        FAILV(VX3H_EUNSUPPORTED, H.collisions ? "simulation with collisions too large for one workgroup of the resident kernel"
                                              : "simulation too large for the resident kernel");
    if (shape < 0 && NV > 0 && b->cfg.path == VX3H_PATH_RESIDENT) {   // a forced resident path takes only what fits
        if (!resOk && (int)H.vmats.size() > RES_MAX_VMATS) FAILV(VX3H_EUNSUPPORTED, "too many voxel materials for the resident kernel");
        if (!resOk && (int)H.lmats.size() > RES_MAX_LMATS) FAILV(VX3H_EUNSUPPORTED, "too many link materials for the resident kernel");
        if (!resOk && nPhases > RES_MAX_PHASES) FAILV(VX3H_EUNSUPPORTED, "too many distinct phase offsets for the resident kernel");
    }
    if (voxcad && b->cfg.path == VX3H_PATH_RESIDENT)
        FAILV(VX3H_EUNSUPPORTED, "VoxCAD stop conditions and equilibrium mode run on the stream path");
    if (H.extras() && b->cfg.path == VX3H_PATH_RESIDENT) FAILV(VX3H_EUNSUPPORTED, "cilia and force fields run on the stream path");
    if ((desc->enable_detach || desc->secondary_experiment || desc->enable_signals || attach || closePairs || targets) &&
        b->cfg.path == VX3H_PATH_RESIDENT)
        FAILV(VX3H_EUNSUPPORTED,
              "detach, attach, secondary experiments, signals, target contacts and close pairs run on the stream path");
    // control block
    SimCtl c;
    memset(&c, 0, sizeof(c));
    c.t = 0;
    c.optDt = 0;
    c.tempAmp = H.tempAmp;
    c.tempPeriod = H.tempPeriod;
    c.dtFrac = H.dtFrac;
    for (int k = 0; k < 3; k++) { c.com[k] = H.com0[k]; c.com0[k] = H.com0[k]; }
    c.voxSize = H.voxSize;
    c.dt = H.dtFixed > 0 ? (float)H.dtFixed : 0.0f;
    c.tInit = H.tInit;
    c.state = VX3H_SIM_RUNNING;
    c.maxSteps = H.maxSteps;
    c.flags = H.flags | (shape >= 0 || tiles ? SC_RESIDENT : 0);
    c.v0 = v0; c.nv = NV; c.l0 = l0; c.nl = NL;
    c.stopOff = tok0;
    c.mat0 = mat0;
    c.nvm = (int)vmIdx.size();
    c.nlm = (int)lmIdx.size();
    c.stopLen = (int)H.stopTok.size();
    c.scType = desc->stop_condition_type;
    c.scValue = desc->stop_condition_value;
    c.eqMode = desc->equilibrium_mode != 0;
    c.dtFixed = H.dtFixed;
    c.histHead = 0;
    for (int k = 0; k < 3; k++) { c.ffOff[k] = ffOff[k]; c.ffLen[k] = (int)H.ffTok[k].size(); }
    if (H.extras()) { c.flags |= SC_EXTRAS; b->anyExtras = true; }
    if (!H.cilia.empty()) c.flags |= SC_CILIA;
    if (desc->enable_detach) { c.flags |= SC_DETACH; b->anyDetach = true; }
    if (targets) {
        c.flags |= SC_TARGETS | SC_EXTRAS;
        b->anyExtras = true;
    }
    int at0 = 0, ap0 = 0, apN = 0;
    if (attach) {   // scratch of the attach pass: 2 ints per voxel, and a pair table that grows on demand
        apN = 2 * NV + 64;
        vx3h_status st;
        if ((st = poolAlloc(b, b->attPool, 2 * NV, at0)) || (st = poolAlloc(b, b->pairPool, apN, ap0))) return st;
        c.at0 = at0;
        c.ap0 = ap0;
        c.apCap = apN;
    }
    if (attach) {
        c.flags |= SC_ATTACH | SC_EXTRAS;
        b->anyAttach = b->anyExtras = true;
        b->anyDetach = true;   // the step-start pass refreshes surfaces after attachments
        c.safety = desc->safety_guard;
        for (int k = 0; k < 5; k++) { c.atOff[k] = atOff[k]; c.atLen[k] = (int)H.atTok[k].size(); }
    }
    c.nlMax = NLcap;
    const bool topoCol = H.collisions && (desc->enable_detach || desc->secondary_experiment || attach);
    if ((desc->enable_signals || attach || topoCol) && NV) {   // neighbour per link direction
        std::vector<int> nb(NV);
        for (int k = 0; k < 6; k++) {
            for (int i = 0; i < NV; i++) {
                const int l = H.voxelLinks[6 * (size_t)i + k];
                nb[i] = l < 0 ? -1 : v0 + (H.lneg[l] == i ? H.lpos[l] : H.lneg[l]);
            }
            UPLOAD(d.nbr + (size_t)k * d.Vcap + v0, nb.data(), NV * 4);
        }
    }
    if (desc->enable_signals && NV) {
        c.flags |= SC_SIGNALS | SC_EXTRAS;
        b->anyExtras = b->anySignals = true;
        const std::vector<double> zero(NV, 0.0);
        for (int k = 0; k < 8; k++) UPLOAD(d.sg + (size_t)k * d.Vcap + v0, zero.data(), NV * 8);
        std::vector<double> col(NV);
        for (int k = 0; k < 5; k++) {
            for (int i = 0; i < NV; i++) {
                const vx3h_material_desc& md = desc->materials[desc->material[i]];
                col[i] = k == 0 ? (double)(md.is_pacemaker != 0) : k == 1 ? md.pacemaker_period : k == 2 ? md.signal_value_decay
                         : k == 3 ? md.signal_time_delay : md.inactive_period;
            }
            UPLOAD(d.sp + (size_t)k * d.Vcap + v0, col.data(), NV * 8);
        }
    }
    if (desc->secondary_experiment) {
        c.flags |= SC_SECONDARY;
        b->anyDetach = true;   // the step-start pass refreshes surfaces after removals too
        c.reinitAfter = desc->reinit_after;
        c.reinitDone = 0;
        std::vector<double> col(NV);
        for (int k = 0; k < 3; k++) {   // saveInitialPosition at creation: the lattice positions
            for (int i = 0; i < NV; i++) col[i] = H.ijk[3 * i + k] * H.voxSize;
            if (NV) UPLOAD(d.ip + (size_t)k * d.Vcap + v0, col.data(), NV * 8);
        }
        for (int i = 0; i < NV; i++) col[i] = desc->materials[desc->material[i]].remove_after;
        if (NV) UPLOAD(d.rt + v0, col.data(), NV * 8);
    }
    if (voxcad) {
        const std::vector<double> unset(3 * VX3_HIST, -1.0);   // ClearHistories
        UPLOAD(b->db.hist + (size_t)s * 3 * VX3_HIST, unset.data(), unset.size() * sizeof(double));
    }
    int tg0 = 0, nTg = 0;
    if (closePairs && NV > 0) {   // registerTargets (VX3_VoxelyzeKernel.cu:536-543): target voxels in voxel order
        std::vector<int> tg;
        for (int i = 0; i < NV; i++)
            if (desc->materials[desc->material[i]].is_target) tg.push_back(v0 + i);
        nTg = (int)tg.size();
        if (nTg) {
            vx3h_status st = poolAlloc(b, b->tgtPool, nTg, tg0);
            if (st) return st;
            UPLOAD(b->db.tgtList + tg0, tg.data(), (size_t)nTg * 4);
        }
        c.closeR = desc->max_dist_close_pairs * H.voxSize;   // MaxDistInVoxelLengthsToCountAsPair * voxSize
        c.tg0 = tg0;
        c.nTg = nTg;
    }
    if (NV == 0) c.state = VX3H_SIM_DONE;   // "No links and no voxels" (VX3_SimulationManager.cu:15-19)
    else if (!H.stopTok.empty() &&
             evalMathTree(H.stopTok.data(), (int)H.stopTok.size(), c.com[0], c.com[1], c.com[2], 0, 0.0, 0, 0, 0, NV) > 0)
        c.state = VX3H_SIM_DONE;
    // collisions: thresholds of CVoxelyze::updateCollisions (Voxelyze.cpp:673-676, float arithmetic),
    // nearby lists of CVX_Voxel::generateNearby (Voxelyze.cpp:394-417) and device pool ranges
    int nb0 = 0, nbN = 0, h0 = 0, hN = 0, near0 = 0, nearN = 0, c0 = 0;
    if (H.collisions && NV > 0) {
        const float br = (float)H.boundingRadius, wd = (float)H.watchDistance;
        const float wrv = 2 * br + wd;
        const float wrMm = (float)(H.voxSize * wrv);
        const float recalc = (float)(H.voxSize * wd / 2);
        c.colThreshSq = wrMm * wrMm;
        c.colRecalcSq = recalc * recalc;
        c.colCell = wrMm;
        c.colInterval = b->cfg.collision_rebuild_interval;
        c.colNearHops = (int)(wrv * 2);
        const int hops = c.colNearHops;
        std::vector<int> nearStart(NV, 0), nearCount(NV, 0), nearList, dist(NV, -1), q, tmp;
        int nSurf = 0;
        for (int v = 0; v < NV; v++) {
            if (H.vflags[v] & VF_INTERIOR) continue;
            nSurf++;
            q.assign(1, v);
            dist[v] = 0;
            for (size_t qi = 0; qi < q.size(); qi++) {
                const int u = q[qi];
                if (dist[u] >= hops) continue;
                for (int d = 0; d < 6; d++) {
                    const int l = H.voxelLinks[6 * u + d];
                    if (l < 0) continue;
                    const int w = H.lneg[l] == u ? H.lpos[l] : H.lneg[l];
                    if (dist[w] < 0) { dist[w] = dist[u] + 1; q.push_back(w); }
                }
            }
            tmp.clear();
            for (int u : q) {
                if (u != v && !(H.vflags[u] & VF_INTERIOR)) tmp.push_back(u);
                dist[u] = -1;
            }
            std::sort(tmp.begin(), tmp.end());
            nearStart[v] = (int)nearList.size();
            nearCount[v] = (int)tmp.size();
            nearList.insert(nearList.end(), tmp.begin(), tmp.end());
        }
        nbN = b->cfg.collision_table_entries > 0 ? b->cfg.collision_table_entries : std::max(64, 64 * nSurf);
        int hb = 10;   // whole blocks of VX3_COL_HBLOCK buckets, at least two buckets per surface voxel
        while ((1 << hb) < 2 * nSurf) hb++;
        hN = 1 << hb;
        nearN = std::max(1, (int)nearList.size());
        if (topoCol) nearN = std::max(nearN + nearN / 2, 64);   // lists change with the topology
        vx3h_status st;
        int hblk0 = 0;
        if ((st = poolAlloc(b, b->nbPool, nbN, nb0)) || (st = poolAlloc(b, b->hashPool, hN / VX3_COL_HBLOCK, hblk0)) ||
            (st = poolAlloc(b, b->nearPool, nearN, near0)) || (st = poolAlloc(b, b->cellPool, NV, c0)))
            return st;
        h0 = hblk0 * VX3_COL_HBLOCK;
        // nearby sets as lattice-offset masks too, where every member is within the mask's
        // offsets: a hop radius of at most 5 and only lattice links (attach makes others)
        if (hops <= VX3_NEAR_RADIUS && !attach) {
            if ((st = uploadNearMasks(b, v0, NV, H.ijk.data(), nearStart, nearCount, nearList))) return st;
            c.flags |= SC_NEARMASK;
        }
        for (int& x : nearStart) x += near0;
        if (!nearList.empty()) UPLOAD(b->db.nearList + near0, nearList.data(), nearList.size() * 4);
        UPLOAD(b->db.nearStart + v0, nearStart.data(), NV * 4);
        UPLOAD(b->db.nearCount + v0, nearCount.data(), NV * 4);
        FILL(b->db.nbCount + v0, 0, NV * 4);
        FILL(b->db.nbStart + v0, 0, NV * 4);
        std::vector<int> own(hN, s);
        UPLOAD(b->db.hashSim + h0, own.data(), hN * 4);
        c.nb0 = nb0;
        c.nbCap = nbN;
        c.h0 = h0;
        c.hBits = hb;
        c.c0 = c0;
        c.near0 = near0;
        c.nearCap = nearN;
        if (topoCol) {   // nearby lists regenerated on the device after topology changes
            c.flags |= SC_TOPOCOL;
            b->anyTopoCol = true;
        }
        c.flags |= SC_REBUILD;
        b->anyColl = true;
        b->db.Ccap = b->nbPool.cap;
        b->db.hashTot = b->hashPool.ra.hi * VX3_COL_HBLOCK;
        b->db.cellHi = b->cellPool.ra.hi;
    }
    UPLOAD(b->db.sims + s, &c, sizeof(c));

    SlotHost& sh = b->slots[s];
    sh = SlotHost();
    sh.used = true;
    sh.v0 = v0; sh.nv = NV; sh.l0 = l0; sh.nl = NL;
    sh.nlCap = NLcap; sh.nlSeen = NL;
    sh.ijk = H.ijk;
    sh.voxSize = H.voxSize;
    for (int k = 0; k < 3; k++) sh.com0[k] = H.com0[k];
    sh.fit = H.fitTok;
    sh.bytes = need;
    sh.shape = shape;
    sh.axisLinks = std::max(perAxis[0], std::max(perAxis[1], perAxis[2]));
    sh.nb0 = nb0; sh.nbN = nbN; sh.h0 = h0; sh.hN = hN; sh.near0 = near0; sh.nearN = nearN; sh.c0 = c0;
    sh.tg0 = tg0; sh.nTg = nTg;
    sh.at0 = at0; sh.ap0 = ap0; sh.apN = apN;
    if (tiles) {   // the tile tables into their pools, once
        vx3h_status st;
        sh.tvN = (int)tplan.vox.size();
        sh.tlN = std::max(1, (int)tplan.link.size());
        sh.txN = std::max(1, tplan.nPub);
        if ((st = poolAlloc(b, b->tileSlotPool, sh.tvN, sh.tv0)) || (st = poolAlloc(b, b->tileLinkPool, sh.tlN, sh.tl0)) ||
            (st = poolAlloc(b, b->tileXchPool, sh.txN, sh.tx0)) || (st = poolAlloc(b, b->tilePartPool, tiles, sh.tp0)))
            return st;
        UPLOAD(b->dTileVox + sh.tv0, tplan.vox.data(), tplan.vox.size() * 4);
        UPLOAD(b->dTileXidx + sh.tv0, tplan.xidx.data(), tplan.xidx.size() * 4);
        if (!tplan.link.empty()) {
            UPLOAD(b->dTileLink + sh.tl0, tplan.link.data(), tplan.link.size() * 4);
            UPLOAD(b->dTileEnds + sh.tl0, tplan.ends.data(), tplan.ends.size() * 4);
        }
        sh.tiles = tiles;
        sh.tplan.tiles = std::move(tplan.tiles);   // the descriptors stay on the host
        sh.tplan.nPub = tplan.nPub;
        b->tiledSims.push_back(s);
        b->tileDirty = true;
    } else if (shape >= 0) {   // the kernel variant with contacts for a simulation with collisions
        sh.shape = shape + (H.collisions ? (int)b->shapes.size() : 0);
        b->resLists[sh.shape].push_back(s);
        b->resDirty = true;
        b->tileDirty = true;
    }
    else b->nStreamSims++;
    sh.vlinks = H.voxelLinks;
    sh.surface.resize(NV);
    for (int i = 0; i < NV; i++) sh.surface[i] = (H.vflags[i] & VF_INTERIOR) ? 0 : 1;
    sh.lneg = H.lneg;
    sh.lpos = H.lpos;
    sh.lratio.resize(NL);
    sh.lfail.resize(NL);
    for (int i = 0; i < NL; i++) { sh.lratio[i] = H.lmats[H.lmat[i]].strainRatio; sh.lfail[i] = H.lmats[H.lmat[i]].epsilonFail; }
    for (int i = 0; i < desc->n_materials; i++) {
        const vx3h_material_desc& md = desc->materials[i];
        sh.palette.push_back({md.matid, md.color[0], md.color[1], md.color[2], md.color[3]});
    }
    sh.matid.resize(NV);
    sh.measured.resize(NV);
    sh.mass.resize(NV);
    for (int i = 0; i < NV; i++) {
        sh.matid[i] = desc->materials[H.vmat[i]].matid;
        sh.measured[i] = H.vmats[H.vmat[i]].isMeasured;
        sh.mass[i] = H.vmats[H.vmat[i]].mass;
    }
    b->bytesUsed += need;
    d.vHi = b->va.hi;
    d.lHi = b->la.hi;
    d.nSims = (int)b->slots.size();
    b->layout++;
    b->dtLo = std::min(b->dtLo, s);
    b->dtHi = std::max(b->dtHi, s + 1);
    *slot = s;
    return VX3H_OK;
}
#undef UPLOAD
#undef FILL
}  // namespace vx3

extern "C" vx3h_status vx3h_batch_step(vx3h_batch* b, int n_steps) {
    if (!b || n_steps < 0) FAILV(VX3H_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(b->device));
    if (b->slots.empty() || n_steps == 0) return VX3H_OK;
    if (b->dtHi > b->dtLo) {
        if (b->realSize == 4) launchDt(b->db, b->stream, b->dtLo, b->dtHi);
        else launchDt(*reinterpret_cast<DevBatch<double>*>(&b->db), b->stream, b->dtLo, b->dtHi);
        HIPCHK(hipGetLastError());
        b->dtLo = 1 << 30;
        b->dtHi = -1;
    }
    const int chunk = b->cfg.steps_per_launch > 0 ? b->cfg.steps_per_launch : 1000;
    if (vx3h_status ps = placeLoads(b, false)) return ps;
    b->resumePending = false;   // every waiting simulation resumes in the first step below
    for (int done = 0; done < n_steps;) {
        const int c = std::min(chunk, n_steps - done);
        vx3h_status st = launchResidentAll(b, c, nullptr);
        if (!st && b->nStreamSims > 0) {
            if (b->cfg.use_graph) {
                const int G = std::min(c, 32);
                st = ensureGraph(b, G);
                for (int r = 0; !st && r < c / G; r++)
                    if (hipGraphLaunch(b->graphExec, b->stream) != hipSuccess) { g_err = "graph launch failed"; st = VX3H_EDEVICE; }
                if (!st && c % G) st = launchSteps(b, c % G);
            } else st = launchSteps(b, c);
        }
        // the main stream waits for the resident launches, also after a failure
        const vx3h_status js = joinResident(b);
        if (st) return st;
        if (js) return js;
        done += c;
    }
    return VX3H_OK;
}

namespace {
StepOpts stepOpts(const vx3h_batch* b) {
    StepOpts o;
    o.contacts = b->anyColl;
    o.extras = b->anyExtras;
    o.prestep = b->anyDetach;
    o.signals = b->anySignals;
    o.attach = b->anyAttach;
    o.topo = b->anyTopoCol;
    o.block = b->cfg.block_size > 0 && b->cfg.block_size <= 256 ? b->cfg.block_size : 256;
    return o;
}
}  // namespace

// ---- the entry point that allocates nothing (vx3h_batch_step_noalloc)
//
// The per-step kernels' one scratch buffer, the link end loads of a step (36 values per
// voxel in the state precision), is a file-scope device array.  Every GPU the library
// runs on holds its own copy until the process ends; host code uses the address
// hipGetSymbolAddress returns for the current device, never the array's name.
static __device__ __align__(256) unsigned char scratch[VX3H_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratchMu;
void* g_scratchPtr[128] = {};

// The device address of `scratch` on `device` (current), looked up once per device.
vx3h_status noallocScratch(int device, void** out) {
    if (device < 0 || device >= 128) FAILV(VX3H_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) FAILV(VX3H_EDEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratchPtr[device] = p;
    }
    *out = g_scratchPtr[device];
    return VX3H_OK;
}

// Before a stepping entry point launches its steps with the end loads in the batch's
// buffer (scratch false) or in the static array (true): a simulation resuming after its
// link pass finds the loads that pass left, copied over if the last step used the other
// buffer.
vx3h_status placeLoads(vx3h_batch* b, bool scratch) {
    const size_t fs = scratch ? (size_t)b->db.vHi : (size_t)b->db.Fs;   // the stride the coming steps use
    const int nv = std::min(b->loadsNv, b->db.vHi);
    if (b->resumePending && scratch != b->loadsInScratch && nv > 0) {
        void* scr = nullptr;
        vx3h_status st = noallocScratch(b->device, &scr);
        if (st) return st;
        const void* src = scratch ? (const void*)b->db.F : scr;
        void* dst = scratch ? scr : (void*)b->db.F;
        // the batch's buffer keeps its contents at its current stride when it grows
        const size_t srcFs = scratch ? (size_t)b->db.Fs : b->loadsFs;
        if (b->realSize == 4) launchCopyLoads(b->db, b->stream, src, srcFs, dst, fs, nv);
        else launchCopyLoads(*reinterpret_cast<DevBatch<double>*>(&b->db), b->stream, src, srcFs, dst, fs, nv);
        HIPCHK(hipGetLastError());
    }
    b->loadsInScratch = scratch;
    b->loadsFs = fs;
    b->loadsNv = b->db.vHi;
    return VX3H_OK;
}
}  // namespace

extern "C" vx3h_status vx3h_batch_step_noalloc(vx3h_batch* b, int n_steps) {
    if (!b || n_steps < 0) FAILV(VX3H_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(b->device));
    void* scr = nullptr;
    vx3h_status st = noallocScratch(b->device, &scr);
    if (st) return st;
    const size_t need = (size_t)36 * (size_t)b->db.vHi * b->realSize;
    if (need > (size_t)VX3H_NOALLOC_SCRATCH_BYTES)
        FAILV(VX3H_EINVAL, "batch too large for the no-allocation step: its end loads need " + std::to_string(need) +
                               " bytes of the " + std::to_string((size_t)VX3H_NOALLOC_SCRATCH_BYTES) + "-byte scratch");
    if (b->slots.empty() || n_steps == 0 || b->db.vHi == 0) return VX3H_OK;
    if (b->dtHi > b->dtLo) {   // recommended step of newly admitted simulations (a kernel; nothing allocated)
        if (b->realSize == 4) launchDt(b->db, b->stream, b->dtLo, b->dtHi);
        else launchDt(*reinterpret_cast<DevBatch<double>*>(&b->db), b->stream, b->dtLo, b->dtHi);
        HIPCHK(hipGetLastError());
        b->dtLo = 1 << 30;
        b->dtHi = -1;
    }
    if ((st = placeLoads(b, true))) return st;
    b->resumePending = false;   // every waiting simulation resumes in the first step below
    const StepOpts o = stepOpts(b);
    for (int i = 0; i < n_steps; i++) {
        if (b->realSize == 4) {
            DevBatch<float> v = b->db;
            v.solo = -2;
            v.F = (float*)scr;
            v.Fs = (size_t)v.vHi;
            launchStep(v, b->stream, o);
        } else {
            DevBatch<double> v = *reinterpret_cast<DevBatch<double>*>(&b->db);
            v.solo = -2;
            v.F = (double*)scr;
            v.Fs = (size_t)v.vHi;
            launchStep(v, b->stream, o);
        }
    }
    HIPCHK(hipGetLastError());
    return VX3H_OK;
}

// VX3_VoxelyzeKernel::doTimeStep(float dt) of one simulation (VX3_VoxelyzeKernel.cu:237-359).
extern "C" vx3h_status vx3h_sim_step(vx3h_batch* b, int slot, float dt, int* ok) {
    if (!b || !ok) FAILV(VX3H_EINVAL, "null argument");
    if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) FAILV(VX3H_ENOTFOUND, "no such slot");
    if (dt != dt) FAILV(VX3H_EINVAL, "dt is NaN");
    HIPCHK(hipSetDevice(b->device));
    *ok = 1;
    if (b->dtHi > b->dtLo) {   // recommended step of newly admitted simulations (the first doTimeStep's)
        if (b->realSize == 4) launchDt(b->db, b->stream, b->dtLo, b->dtHi);
        else launchDt(*reinterpret_cast<DevBatch<double>*>(&b->db), b->stream, b->dtLo, b->dtHi);
        HIPCHK(hipGetLastError());
        b->dtLo = 1 << 30;
        b->dtHi = -1;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    SimCtl c;
    HIPCHK(hipMemcpy(&c, b->db.sims + slot, sizeof(c), hipMemcpyDeviceToHost));
    if (c.state == VX3H_SIM_DIVERGED) { *ok = 0; return VX3H_OK; }
    if (c.state == VX3H_SIM_EMPTY || c.state == VX3H_SIM_FAILED || b->slots[slot].nv == 0) return VX3H_OK;
    if (c.flags & SC_COL_GROW) {   // a collision table to enlarge first (vx3h_batch_poll does it)
        int n = 0;
        vx3h_status st = vx3h_batch_poll(b, nullptr, 0, &n);
        if (st) return st;
        HIPCHK(hipMemcpy(&c, b->db.sims + slot, sizeof(c), hipMemcpyDeviceToHost));
    }
    if (dt == 0) {   // CurStepCount++ and return: the temperature for the unchanged time is already in place
        c.steps++;
        if (c.maxSteps > 0 && c.steps >= c.maxSteps) c.state = VX3H_SIM_DONE;
        HIPCHK(hipMemcpy(b->db.sims + slot, &c, sizeof(c), hipMemcpyHostToDevice));
        return VX3H_OK;
    }
    // dt < 0: DtFrac * OptimalDt (or the fixed step), the simulation's own dt
    const SimCtl saved = c;
    const float stepDt = dt < 0 ? c.dt : dt;
    c.dt = stepDt;
    // CycleStep = int(TempPeriod / dt) of this call's dt (VX3_VoxelyzeKernel.cu:314-316)
    c.cycleStep = c.tempPeriod / (double)stepDt >= 1.0 ? (int)(c.tempPeriod / (double)stepDt) : 0;
    c.flags &= ~SC_RESIDENT;   // the per-step kernels take it, whichever path it runs on
    if (c.state == VX3H_SIM_DONE) c.state = VX3H_SIM_RUNNING;   // stepping past the stop condition is the caller's call
    HIPCHK(hipMemcpy(b->db.sims + slot, &c, sizeof(c), hipMemcpyHostToDevice));
    if (vx3h_status ps = placeLoads(b, false)) return ps;
    StepOpts o = stepOpts(b);
    for (int attempt = 0; attempt < 8; attempt++) {
        if (b->realSize == 4) {
            DevBatch<float> v = b->db;
            v.solo = slot;
            launchStep(v, b->stream, o);
        } else {
            DevBatch<double> v = *reinterpret_cast<DevBatch<double>*>(&b->db);
            v.solo = slot;
            launchStep(v, b->stream, o);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(b->stream));
        HIPCHK(hipMemcpy(&c, b->db.sims + slot, sizeof(c), hipMemcpyDeviceToHost));
        if (!(c.flags & SC_COL_GROW)) break;
        int n = 0;   // the collision table was too small: enlarge it, then take the step again
        vx3h_status st = vx3h_batch_poll(b, nullptr, 0, &n);
        if (st) return st;
    }
    if (c.flags & SC_COL_GROW) FAILV(VX3H_EDEVICE, "collision table did not grow");
    // the simulation's own step size and sampling period stay for the next call
    c.dt = saved.dt;
    c.cycleStep = saved.cycleStep;
    c.flags = (c.flags & ~SC_RESIDENT) | (saved.flags & SC_RESIDENT);
    HIPCHK(hipMemcpy(b->db.sims + slot, &c, sizeof(c), hipMemcpyHostToDevice));
    *ok = c.state == VX3H_SIM_DIVERGED ? 0 : 1;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_batch_kernel_times(vx3h_batch* b, int n_steps, double* avg_ms) {
    if (!b || !avg_ms || n_steps <= 0) FAILV(VX3H_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(b->device));
    for (int k = 0; k < VX3H_N_PASSES; k++) avg_ms[k] = 0;
    if (b->slots.empty()) return VX3H_OK;
    std::vector<hipEvent_t> ev(2 * PASS_COUNT * n_steps, nullptr);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    const StepOpts o = stepOpts(b);
    for (int i = 0; i < n_steps; i++) {
        hipEvent_t* e = &ev[2 * PASS_COUNT * i];
        if (b->realSize == 4) launchStep(b->db, b->stream, o, e);
        else launchStep(*reinterpret_cast<DevBatch<double>*>(&b->db), b->stream, o, e);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    const bool ran[PASS_COUNT] = {false, b->db.lHi > 0, o.contacts, true, true};   // Poisson: part of the voxel pass
    for (int k = 0; k < PASS_COUNT; k++) {
        if (!ran[k] || b->nStreamSims == 0) continue;
        double tot = 0;
        for (int i = 0; i < n_steps; i++) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev[2 * PASS_COUNT * i + 2 * k], ev[2 * PASS_COUNT * i + 2 * k + 1]));
            tot += ms;
        }
        avg_ms[k] = tot / n_steps;
    }
    for (auto& e : ev) hipEventDestroy(e);
    // resident kernels: one launch of n_steps steps per shape in use and per tiled launch
    // group; reports the launch durations' sum
    vx3h_status st = syncTiles(b);
    if (st) return st;
    const size_t nl = b->resLists.size() + b->tileGroups.size();
    std::vector<hipEvent_t> rev(2 * nl, nullptr);
    for (auto& e : rev) HIPCHK(hipEventCreate(&e));
    st = launchResidentAll(b, n_steps, rev.data());
    if (st || (st = joinResident(b))) return st;
    HIPCHK(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < nl; i++) {
        if (i < b->resLists.size() && (b->resLists[i].empty() || (b->absorbed && i < b->shapes.size()))) continue;
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, rev[2 * i], rev[2 * i + 1]));
        avg_ms[5] += ms;
    }
    for (auto& e : rev) hipEventDestroy(e);
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_sim_workgroups(vx3h_batch* b, int slot, int* workgroups) {
    if (!b || !workgroups) FAILV(VX3H_EINVAL, "null argument");
    if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) FAILV(VX3H_ENOTFOUND, "no such slot");
    const SlotHost& sh = b->slots[slot];
    *workgroups = sh.tiles ? sh.tiles : sh.shape >= 0 ? 1 : 0;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_batch_set_timing(vx3h_batch* b, int enable) {
    if (!b) FAILV(VX3H_EINVAL, "null batch");
    b->timing = enable != 0;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_batch_timing(vx3h_batch* b, double* total_ms, int* launches, long long* steps) {
    if (!b || !total_ms || !launches || !steps) FAILV(VX3H_EINVAL, "null argument");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    double tot = 0;
    for (auto& pr : b->evUsed) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
        tot += ms;
        b->evPool.push_back(pr.first);
        b->evPool.push_back(pr.second);
    }
    *total_ms = tot;
    *launches = (int)b->evUsed.size();
    *steps = b->timedSteps;
    b->evUsed.clear();
    b->timedSteps = 0;
    return VX3H_OK;
}

// Plain streaming copy: the reference point for HBM traffic counters and for the
// bandwidth a simple kernel reaches on this device.
__global__ void __launch_bounds__(256) k_calib_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

extern "C" vx3h_status vx3h_calib_copy(int device, size_t bytes, int reps, double* ms_per_copy) {
    if (bytes < 16 || reps <= 0 || !ms_per_copy) FAILV(VX3H_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(device));
    const size_t n = bytes / 16;
    void *a = nullptr, *c = nullptr;
    HIPCHK(hipMalloc(&a, n * 16));
    if (hipMalloc(&c, n * 16) != hipSuccess) { hipFree(a); FAILV(VX3H_ENOMEM, "calibration buffers"); }
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipMemsetAsync(a, 0, n * 16, st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int grid = 256 * 8 * 4;
    hipLaunchKernelGGL(k_calib_copy, dim3(grid), dim3(256), 0, st, (const float4*)a, (float4*)c, n);   // warm
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_calib_copy, dim3(grid), dim3(256), 0, st, (const float4*)a, (float4*)c, n);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_copy = ms / reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(st);
    hipFree(a);
    hipFree(c);
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_batch_sync(vx3h_batch* b) {
    if (!b) FAILV(VX3H_EINVAL, "null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    return VX3H_OK;
}

namespace vx3 {
vx3h_status pollOne(vx3h_batch* b, int s, const SimCtl& c, vx3h_sim_status& o, std::vector<double>* posOut) {
    const SlotHost& sh = b->slots[s];
    memset(&o, 0, sizeof(o));
    o.slot = s;
    o.state = c.state;
    o.steps = c.steps;
    o.num_voxel = sh.nv;
    o.current_time = c.t;
    o.dt = c.dt;
    o.recommended_dt = c.optDt;
    o.collision_count = c.hits;
    o.n_links = c.nl;
    o.recent_angle = c.recentAngle;
    // initial center of mass and positions: the admission's, or a secondary experiment's reset
    for (int k = 0; k < 3; k++) o.initial_com[k] = c.com0[k];
    std::vector<double> x, y, z;
    vx3h_status st;
    const size_t rs = b->realSize;
    if ((st = readReal(b, at(b->db.px, rs, sh.v0), sh.nv, x))) return st;
    if ((st = readReal(b, at(b->db.py, rs, sh.v0), sh.nv, y))) return st;
    if ((st = readReal(b, at(b->db.pz, rs, sh.v0), sh.nv, z))) return st;
    std::vector<double> ip;
    if (c.flags & SC_SECONDARY) {
        ip.resize(3 * (size_t)sh.nv);
        for (int k = 0; k < 3; k++)
            HIPCHK(hipMemcpy(ip.data() + (size_t)k * sh.nv, b->db.ip + (size_t)k * b->db.Vcap + sh.v0, sh.nv * 8, hipMemcpyDeviceToHost));
    }
    // updateCurrentCenterOfMass (VX3_VoxelyzeKernel.cu:477-493) + collectResults totals (VX3_SimulationManager.cu:442-452)
    double tot = 0, sx = 0, sy = 0, sz = 0, dist = 0;
    int nm = 0;
    for (int i = 0; i < sh.nv; i++) {
        if (!sh.measured[i]) continue;
        double m = sh.mass[i];
        sx += x[i] * m; sy += y[i] * m; sz += z[i] * m;
        tot += m;
        nm++;
        const double ix = ip.empty() ? sh.ijk[3 * i] * sh.voxSize : ip[i];
        const double iy = ip.empty() ? sh.ijk[3 * i + 1] * sh.voxSize : ip[sh.nv + i];
        const double iz = ip.empty() ? sh.ijk[3 * i + 2] * sh.voxSize : ip[2 * (size_t)sh.nv + i];
        double dx = x[i] - ix, dy = y[i] - iy, dz = z[i] - iz;
        dist += sqrt(dx * dx + dy * dy + dz * dz);
    }
    if (tot != 0) { double inv = 1.0 / tot; o.current_com[0] = inv * sx; o.current_com[1] = inv * sy; o.current_com[2] = inv * sz; }
    o.num_measured_voxel = nm;
    o.total_distance_of_all_voxels = dist;
    double off[3] = {o.current_com[0] - c.com0[0], o.current_com[1] - c.com0[1], o.current_com[2] - c.com0[2]};
    // computeFitness (VX3_VoxelyzeKernel.cu:530-534)
    o.num_close_pairs = c.closePairs;
    o.target_closeness = c.closeness;
    o.fitness = evalMathTree(sh.fit.data(), (int)sh.fit.size(), off[0], off[1], off[2], c.hits, c.t, c.recentAngle, c.closeness,
                             c.closePairs, sh.nv);
    if (posOut) {
        posOut->resize(3 * sh.nv);
        for (int i = 0; i < sh.nv; i++) { (*posOut)[3 * i] = x[i]; (*posOut)[3 * i + 1] = y[i]; (*posOut)[3 * i + 2] = z[i]; }
    }
    return VX3H_OK;
}
}  // namespace vx3

namespace {
// A fresh range of `need` entries for one simulation in a pool (the old range is released).
vx3h_status regrow(vx3h_batch* b, DevPool& p, int& start, int& len, int need) {
    p.ra.release(start, len);
    vx3h_status st = poolAlloc(b, p, need, start);
    if (st) return st;
    len = need;
    return VX3H_OK;
}

// The nearby lists of one simulation from its current links (the device search did not
// fit in its queue): generateNearby (Voxelyze.cpp:394-417) on the host, uploaded.
vx3h_status hostNearby(vx3h_batch* b, int s, SimCtl& c) {
    SlotHost& sh = b->slots[s];
    const int NV = sh.nv;
    const size_t V = (size_t)b->db.Vcap;
    std::vector<int> fl(NV), nbr(6 * (size_t)NV);
    HIPCHK(hipMemcpy(fl.data(), b->db.vflags + sh.v0, NV * 4, hipMemcpyDeviceToHost));
    for (int d = 0; d < 6; d++) HIPCHK(hipMemcpy(nbr.data() + (size_t)d * NV, b->db.nbr + d * V + sh.v0, NV * 4, hipMemcpyDeviceToHost));
    std::vector<int> start(NV, 0), count(NV, 0), list, mark(NV, -1), q;
    for (int v = 0; v < NV; v++) {
        if (fl[v] & VF_INTERIOR) continue;
        q.assign(1, v);
        mark[v] = v;
        size_t cur = 0;
        for (int k = 0; k < c.colNearHops; k++) {
            const size_t end = q.size();
            while (cur < end) {
                const int u = q[cur++];
                for (int d = 0; d < 6; d++) {
                    if (!(fl[u] & (1 << (VF_LINKMASK_SHIFT + d)))) continue;
                    const int w = nbr[(size_t)d * NV + u] - sh.v0;
                    if (w < 0 || w >= NV || mark[w] == v) continue;
                    mark[w] = v;
                    q.push_back(w);
                }
            }
        }
        std::vector<int> t;
        for (size_t i = 1; i < q.size(); i++) if (!(fl[q[i]] & VF_INTERIOR)) t.push_back(q[i]);
        std::sort(t.begin(), t.end());
        start[v] = (int)list.size();
        count[v] = (int)t.size();
        list.insert(list.end(), t.begin(), t.end());
    }
    if ((int)list.size() > sh.nearN) {
        vx3h_status st = regrow(b, b->nearPool, sh.near0, sh.nearN, (int)list.size() * 3 / 2 + 64);
        if (st) return st;
        c.near0 = sh.near0;
        c.nearCap = sh.nearN;
    }
    if (c.flags & SC_NEARMASK) {
        std::vector<int> ijk(3 * (size_t)NV);
        HIPCHK(hipMemcpy(ijk.data(), b->db.vijk + 3 * (size_t)sh.v0, (size_t)NV * 12, hipMemcpyDeviceToHost));
        if (vx3h_status st = uploadNearMasks(b, sh.v0, NV, ijk.data(), start, count, list)) return st;
    }
    for (int& x : start) x += sh.near0;
    if (!list.empty()) HIPCHK(hipMemcpy(b->db.nearList + sh.near0, list.data(), list.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->db.nearStart + sh.v0, start.data(), NV * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->db.nearCount + sh.v0, count.data(), NV * 4, hipMemcpyHostToDevice));
    return VX3H_OK;
}

// Simulations whose collision tables proved too small stand still (SC_COL_GROW: no pass
// has changed their state since; with SC_LINKS_DONE, since the link pass of the step
// that stopped, whose end loads are kept).  A neighbour or nearby table gets room for
// half as many entries again as the rebuild needed (at least twice the old one); nearby
// lists whose search outgrew the device queue are built here.  The simulation then
// rebuilds at its next step, which runs (or, after its link pass, resumes) as it would
// have.  The batch's streams are idle.
vx3h_status growCollisionTables(vx3h_batch* b, std::vector<SimCtl>& ctl) {
    for (int s = 0; s < (int)ctl.size(); s++) {
        if (!b->slots[s].used || !(ctl[s].flags & SC_COL_GROW)) continue;
        SlotHost& sh = b->slots[s];
        SimCtl& c = ctl[s];
        vx3h_status st;
        if (sh.shape >= 0) {   // contact lists beyond the resident kernel's LDS table: the per-step kernels take it
            auto& L = b->resLists[sh.shape];
            L.erase(std::remove(L.begin(), L.end(), s), L.end());
            b->resDirty = true;
            b->tileDirty = true;
            sh.shape = -1;
            b->nStreamSims++;
            c.flags &= ~SC_RESIDENT;
            b->layout++;
        }
        if (c.nbNeed > c.nbCap) {
            if ((st = regrow(b, b->nbPool, sh.nb0, sh.nbN, std::max(2 * sh.nbN, c.nbNeed + c.nbNeed / 2)))) return st;
            c.nb0 = sh.nb0;
            c.nbCap = sh.nbN;
        }
        if (c.nearNeed == (1 << 30)) {
            if ((st = hostNearby(b, s, c))) return st;
            c.flags &= ~SC_NEARDIRTY;
        } else if (c.nearNeed > c.nearCap) {   // the device rebuilds the lists again, in the larger table
            if ((st = regrow(b, b->nearPool, sh.near0, sh.nearN, std::max(2 * sh.nearN, c.nearNeed + c.nearNeed / 2)))) return st;
            c.near0 = sh.near0;
            c.nearCap = sh.nearN;
        }
        if (c.apNeed > c.apCap) {
            if ((st = regrow(b, b->pairPool, sh.ap0, sh.apN, std::max(2 * sh.apN, c.apNeed + c.apNeed / 2)))) return st;
            c.ap0 = sh.ap0;
            c.apCap = sh.apN;
        }
        c.nbNeed = c.nearNeed = c.apNeed = 0;
        c.nearUsed = 0;
        if (c.flags & SC_LINKS_DONE) b->resumePending = true;
        c.flags = (c.flags & ~SC_COL_GROW) | SC_REBUILD;
        HIPCHK(hipMemcpy(b->db.sims + s, &c, sizeof(SimCtl), hipMemcpyHostToDevice));
        b->db.Ccap = b->nbPool.cap;
        b->tableGrowths++;
    }
    return VX3H_OK;
}
}  // namespace

namespace vx3 {
vx3h_status pollFinished(vx3h_batch* b, std::vector<vx3h_sim_status>& finished) {
    finished.clear();
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    std::vector<SimCtl> ctl(b->slots.size());
    if (!ctl.empty()) HIPCHK(hipMemcpy(ctl.data(), b->db.sims, ctl.size() * sizeof(SimCtl), hipMemcpyDeviceToHost));
    if (b->anyColl || b->anyAttach) {
        vx3h_status gs = growCollisionTables(b, ctl);
        if (gs) return gs;
    }
    for (int s = 0; s < (int)b->slots.size(); s++) {
        if (!b->slots[s].used || ctl[s].state == VX3H_SIM_RUNNING) continue;
        finished.emplace_back();
        vx3h_status st = pollOne(b, s, ctl[s], finished.back(), nullptr);
        if (st) return st;
    }
    return VX3H_OK;
}
}  // namespace vx3

extern "C" vx3h_status vx3h_batch_poll(vx3h_batch* b, vx3h_sim_status* out, int cap, int* n) {
    if (!b || !n) FAILV(VX3H_EINVAL, "null argument");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    std::vector<SimCtl> ctl(b->slots.size());
    if (!ctl.empty()) HIPCHK(hipMemcpy(ctl.data(), b->db.sims, ctl.size() * sizeof(SimCtl), hipMemcpyDeviceToHost));
    if (b->anyColl || b->anyAttach) {
        vx3h_status gs = growCollisionTables(b, ctl);
        if (gs) return gs;
    }
    int k = 0;
    for (int s = 0; s < (int)b->slots.size(); s++) {
        if (!b->slots[s].used) continue;
        if (out && k < cap) {
            vx3h_status st = pollOne(b, s, ctl[s], out[k], nullptr);
            if (st) return st;
        }
        k++;
    }
    *n = k;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_sim_retire(vx3h_batch* b, int slot) {
    if (!b) FAILV(VX3H_EINVAL, "null batch");
    if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) FAILV(VX3H_ENOTFOUND, "no such slot");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    SlotHost& sh = b->slots[slot];
    if (sh.nv) HIPCHK(hipMemset(b->db.vsim + sh.v0, 0xFF, (size_t)sh.nv * 4));
    if (sh.nlCap) HIPCHK(hipMemset(b->db.lsim + sh.l0, 0xFF, (size_t)sh.nlCap * 4));
    SimCtl c;
    memset(&c, 0, sizeof(c));
    c.state = VX3H_SIM_EMPTY;
    HIPCHK(hipMemcpy(b->db.sims + slot, &c, sizeof(c), hipMemcpyHostToDevice));
    b->va.release(sh.v0, sh.nv);
    b->la.release(sh.l0, sh.nlCap);
    if (sh.hN) {
        HIPCHK(hipMemset(b->db.hashSim + sh.h0, 0xFF, (size_t)sh.hN * 4));
        b->hashPool.ra.release(sh.h0 / VX3_COL_HBLOCK, sh.hN / VX3_COL_HBLOCK);
        b->nbPool.ra.release(sh.nb0, sh.nbN);
        b->nearPool.ra.release(sh.near0, sh.nearN);
        b->cellPool.ra.release(sh.c0, sh.nv);
        b->db.hashTot = b->hashPool.ra.hi * VX3_COL_HBLOCK;
        b->db.cellHi = b->cellPool.ra.hi;
    }
    if (sh.nTg) b->tgtPool.ra.release(sh.tg0, sh.nTg);
    if (sh.apN) {
        b->attPool.ra.release(sh.at0, 2 * sh.nv);
        b->pairPool.ra.release(sh.ap0, sh.apN);
    }
    if (sh.tiles) {
        b->tiledSims.erase(std::remove(b->tiledSims.begin(), b->tiledSims.end(), slot), b->tiledSims.end());
        b->tileSlotPool.ra.release(sh.tv0, sh.tvN);
        b->tileLinkPool.ra.release(sh.tl0, sh.tlN);
        b->tileXchPool.ra.release(sh.tx0, sh.txN);
        b->tilePartPool.ra.release(sh.tp0, sh.tiles);
        b->tileDirty = true;
    } else if (sh.shape >= 0) {
        auto& L = b->resLists[sh.shape];
        L.erase(std::remove(L.begin(), L.end(), slot), L.end());
        b->resDirty = true;
        b->tileDirty = true;
    } else b->nStreamSims--;
    if (sh.oneTile) {
        b->tileSlotPool.ra.release(sh.tv0, sh.tvN);
        b->tileLinkPool.ra.release(sh.tl0, sh.tlN);
        b->tileXchPool.ra.release(sh.tx0, sh.txN);
        b->tilePartPool.ra.release(sh.tp0, 1);
    }
    b->bytesUsed -= sh.bytes;
    sh = SlotHost();
    b->db.vHi = b->va.hi;
    b->db.lHi = b->la.hi;
    b->layout++;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_sim_collision_pairs(vx3h_batch* b, int slot, int* pairs, int cap, int* n) {
    if (!b || !n || cap < 0) FAILV(VX3H_EINVAL, "bad argument");
    if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) FAILV(VX3H_ENOTFOUND, "no such slot");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    const SlotHost& sh = b->slots[slot];
    *n = 0;
    if (!sh.nbN || !sh.nv) return VX3H_OK;
    std::vector<int> st(sh.nv), cnt(sh.nv), fl(sh.nv);
    HIPCHK(hipMemcpy(st.data(), b->db.nbStart + sh.v0, sh.nv * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cnt.data(), b->db.nbCount + sh.v0, sh.nv * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fl.data(), b->db.vflags + sh.v0, sh.nv * 4, hipMemcpyDeviceToHost));
    std::vector<int> list(sh.nbN);
    HIPCHK(hipMemcpy(list.data(), b->db.nbList + sh.nb0, (size_t)sh.nbN * 4, hipMemcpyDeviceToHost));
    int k = 0;
    for (int v = 0; v < sh.nv; v++) {
        if (fl[v] & VF_INTERIOR) continue;
        for (int i = 0; i < cnt[v]; i++) {
            const int idx = st[v] - sh.nb0 + i;
            if (idx < 0 || idx >= sh.nbN) FAILV(VX3H_EDEVICE, "collision list out of range");
            const int u = list[idx] - sh.v0;
            if (u <= v) continue;
            if (pairs && k < cap) { pairs[2 * k] = v; pairs[2 * k + 1] = u; }
            k++;
        }
    }
    *n = k;
    return VX3H_OK;
}

// Brings the host shadows of a simulation's topology up to the links attach has made on
// the device (they follow the admitted links in the order they were made).
static vx3h_status syncAttached(vx3h_batch* b, int slot) {
    SlotHost& sh = b->slots[slot];
    if (sh.nlCap <= sh.nl) return VX3H_OK;
    SimCtl c;
    HIPCHK(hipMemcpy(&c, b->db.sims + slot, sizeof(SimCtl), hipMemcpyDeviceToHost));
    const int n = std::min(c.nl, sh.nlCap) - sh.nlSeen;
    if (n <= 0) return VX3H_OK;
    std::vector<int> ln(n), lp(n), lm(n);
    const size_t o = (size_t)sh.l0 + sh.nlSeen;
    HIPCHK(hipMemcpy(ln.data(), b->db.lneg + o, n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(lp.data(), b->db.lpos + o, n * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(lm.data(), b->db.lmeta + o, n * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        const int li = sh.nlSeen + i, vn = ln[i] - sh.v0, vp = lp[i] - sh.v0, axis = lm[i] & 3;
        const LinkMatDev& m = b->lmats.items[lm[i] >> 2];
        sh.lneg.push_back(vn);
        sh.lpos.push_back(vp);
        sh.lratio.push_back(m.strainRatio);
        sh.lfail.push_back(m.epsilonFail);
        sh.vlinks[6 * (size_t)vn + 2 * axis] = li;
        sh.vlinks[6 * (size_t)vp + 2 * axis + 1] = li;
    }
    sh.nlSeen += n;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_sim_read_state(vx3h_batch* b, int slot, vx3h_state_view* view) {
    if (!b || !view) FAILV(VX3H_EINVAL, "null argument");
    if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) FAILV(VX3H_ENOTFOUND, "no such slot");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (view->link_capacity > 0) { vx3h_status s0 = syncAttached(b, slot); if (s0) return s0; }
    const SlotHost& sh = b->slots[slot];
    // links written: the admitted ones, or with a capacity also those attach made
    const int nlOut = view->link_capacity > 0 ? std::min(view->link_capacity, sh.nlSeen) : sh.nl;
    const DevBatch<float>& d = b->db;
    const size_t rs = b->realSize;
    std::vector<double> t0, t1, t2, t3;
    vx3h_status st;
    auto get3 = [&](void* a, void* bb, void* c, double* out) -> vx3h_status {
        vx3h_status s2;
        if ((s2 = readReal(b, at(a, rs, sh.v0), sh.nv, t0))) return s2;
        if ((s2 = readReal(b, at(bb, rs, sh.v0), sh.nv, t1))) return s2;
        if ((s2 = readReal(b, at(c, rs, sh.v0), sh.nv, t2))) return s2;
        for (int i = 0; i < sh.nv; i++) { out[3 * i] = t0[i]; out[3 * i + 1] = t1[i]; out[3 * i + 2] = t2[i]; }
        return VX3H_OK;
    };
    if (view->pos && (st = get3(d.px, d.py, d.pz, view->pos))) return st;
    if (view->lin_mom && (st = get3(d.mx, d.my, d.mz, view->lin_mom))) return st;
    if (view->ang_mom && (st = get3(d.ax, d.ay, d.az, view->ang_mom))) return st;
    if (view->orient) {
        if ((st = readReal(b, at(d.ow, rs, sh.v0), sh.nv, t3))) return st;
        if ((st = readReal(b, at(d.ox, rs, sh.v0), sh.nv, t0))) return st;
        if ((st = readReal(b, at(d.oy, rs, sh.v0), sh.nv, t1))) return st;
        if ((st = readReal(b, at(d.oz, rs, sh.v0), sh.nv, t2))) return st;
        for (int i = 0; i < sh.nv; i++) {
            view->orient[4 * i] = t3[i]; view->orient[4 * i + 1] = t0[i]; view->orient[4 * i + 2] = t1[i]; view->orient[4 * i + 3] = t2[i];
        }
    }
    if (view->vox_flags) HIPCHK(hipMemcpy(view->vox_flags, d.vflags + sh.v0, sh.nv * 4, hipMemcpyDeviceToHost));
    if (view->init_pos) {
        SimCtl c;
        HIPCHK(hipMemcpy(&c, b->db.sims + slot, sizeof(SimCtl), hipMemcpyDeviceToHost));
        std::vector<double> col(sh.nv);
        for (int k = 0; k < 3; k++) {
            if (c.flags & SC_SECONDARY)
                HIPCHK(hipMemcpy(col.data(), d.ip + (size_t)k * d.Vcap + sh.v0, sh.nv * 8, hipMemcpyDeviceToHost));
            else
                for (int i = 0; i < sh.nv; i++) col[i] = sh.ijk[3 * i + k] * sh.voxSize;
            for (int i = 0; i < sh.nv; i++) view->init_pos[3 * i + k] = col[i];
        }
    }
    if (view->temp) {
        std::vector<float> tf(sh.nv);
        HIPCHK(hipMemcpy(tf.data(), d.tfac + sh.v0, sh.nv * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < sh.nv; i++) view->temp[i] = tf[i];   // thermal size factor 1 + T*CTE
    }
    if (nlOut && (view->link_strain || view->link_max_strain)) {
        std::vector<float> a(nlOut);
        if (view->link_strain) {
            HIPCHK(hipMemcpy(a.data(), d.strain + sh.l0, nlOut * 4, hipMemcpyDeviceToHost));
            for (int i = 0; i < nlOut; i++) view->link_strain[i] = a[i];
        }
        if (view->link_max_strain) {
            HIPCHK(hipMemcpy(a.data(), d.maxStrain + sh.l0, nlOut * 4, hipMemcpyDeviceToHost));
            for (int i = 0; i < nlOut; i++) view->link_max_strain[i] = a[i];
        }
    }
    if (nlOut && view->link_small_angle) {
        std::vector<int> f(nlOut);
        HIPCHK(hipMemcpy(f.data(), d.lflags + sh.l0, nlOut * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < nlOut; i++) view->link_small_angle[i] = (f[i] & LF_SMALL_ANGLE) ? 1 : 0;
    }
    return VX3H_OK;
}

namespace {

void appendf(std::string& out, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void appendf(std::string& out, const char* fmt, ...) {
    char tmp[256];
    va_list ap;
    va_start(ap, fmt);
    const int n = vsnprintf(tmp, sizeof(tmp), fmt, ap);
    va_end(ap);
    if (n < (int)sizeof(tmp)) { out.append(tmp, n > 0 ? n : 0); return; }
    std::vector<char> big(n + 1);
    va_start(ap, fmt);
    vsnprintf(big.data(), big.size(), fmt, ap);
    va_end(ap);
    out.append(big.data(), n);
}

vx3h_status putText(const std::string& t, char* buf, size_t cap, size_t* len) {
    if (len) *len = t.size() + 1;
    if (buf && cap) {
        const size_t k = std::min(cap - 1, t.size());
        memcpy(buf, t.data(), k);
        buf[k] = 0;
    }
    return VX3H_OK;
}

}  // namespace

// VX3_SimulationManager.cu:40-49
extern "C" vx3h_status vx3h_sim_history_header(vx3h_batch* b, int slot, char* buf, size_t cap, size_t* len) {
    if (!b) FAILV(VX3H_EINVAL, "null batch");
    if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) FAILV(VX3H_ENOTFOUND, "no such slot");
    const SlotHost& sh = b->slots[slot];
    std::string t;
    appendf(t, "\n{{{setting}}}<rescale>0.001</rescale>\n");
    for (const auto& m : sh.palette)
        appendf(t, "{{{setting}}}<matcolor><id>%d</id><r>%.2f</r><g>%.2f</g><b>%.2f</b><a>%.2f</a></matcolor>\n", m[0],
                m[1] / 255., m[2] / 255., m[3] / 255., m[4] / 255.);
    appendf(t, "\n{{{setting}}}<voxel_size>%f</voxel_size>\n", sh.voxSize);
    return putText(t, buf, cap, len);
}

// VX3_SimulationManager.cu:70-112 with CVX_Voxel::cornerOffset (VX3_Voxel.cu:147-159) and
// Quat3D::AngleDegrees (Utils/VX3_Quat3D.h:301-306) evaluated on the host from the
// device state
extern "C" vx3h_status vx3h_sim_history_record(vx3h_batch* b, int slot, int step_label, int record_voxel, int record_link,
                                               char* buf, size_t cap, size_t* len) {
    if (!b) FAILV(VX3H_EINVAL, "null batch");
    if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) FAILV(VX3H_ENOTFOUND, "no such slot");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    { vx3h_status s0 = syncAttached(b, slot); if (s0) return s0; }
    const SlotHost& sh = b->slots[slot];
    const int NV = sh.nv, NL = sh.nlSeen;   // the admitted links and those attach made
    std::vector<double> pos(3 * (size_t)NV), ori(4 * (size_t)NV), tf(NV), strain(NL), maxStrain(NL);
    std::vector<int> vflags(NV), lflags(NL);
    vx3h_state_view view{};
    view.pos = pos.data();
    view.orient = ori.data();
    view.temp = tf.data();
    view.vox_flags = vflags.data();
    view.link_strain = strain.data();
    view.link_max_strain = maxStrain.data();
    view.link_capacity = NL;
    vx3h_status st = vx3h_sim_read_state(b, slot, &view);
    if (st) return st;
    if (NL) HIPCHK(hipMemcpy(lflags.data(), b->db.lflags + sh.l0, NL * 4, hipMemcpyDeviceToHost));
    std::vector<double> sig;
    int n = 0;
    if ((st = vx3h_batch_poll(b, nullptr, 0, &n))) return st;
    SimCtl c;
    HIPCHK(hipMemcpy(&c, b->db.sims + slot, sizeof(SimCtl), hipMemcpyDeviceToHost));
    if (c.flags & SC_SIGNALS) {
        sig.resize(NV);
        HIPCHK(hipMemcpy(sig.data(), b->db.sg + sh.v0, NV * 8, hipMemcpyDeviceToHost));
    }
    const double vs = 1 / 0.001;
    std::string t;
    if (record_voxel) {
        appendf(t, "<<<Step%d Time:%f>>>", step_label, c.t);
        for (int v = 0; v < NV; v++) {
            if (vflags[v] & (VF_INTERIOR | VF_REMOVED)) continue;   // surface voxels as of this step's start
            const double* p = &pos[3 * (size_t)v];
            const double* q = &ori[4 * (size_t)v];
            appendf(t, "%.1f,%.1f,%.1f,", p[0] * vs, p[1] * vs, p[2] * vs);
            const double w = q[0] > 1 ? 1 : q[0];
            appendf(t, "%.1f,%.2f,%.2f,%.2f,", 2.0 * acos(w) * 57.29577951308232, q[1], q[2], q[3]);
            // baseSize(): nominal size times the float thermal factor 1 + T*CTE
            const double half = 0.5 * (sh.voxSize * (double)(float)tf[v]);
            double corner[2][3];
            for (int cI = 0; cI < 2; cI++) {   // NNN, then PPP
                const bool posLink = cI == 1;
                for (int i = 0; i < 3; i++) {
                    const int l = sh.vlinks[6 * (size_t)v + 2 * i + (posLink ? 0 : 1)];
                    double sv;
                    if (l >= 0 && !(lflags[l] & LF_GONE) && !(sh.lfail[l] != -1.0f && (float)maxStrain[l] > sh.lfail[l])) {
                        const float s = (float)strain[l], r = sh.lratio[l];
                        const float ax = posLink ? 2.0f * s * r / (1.0f + r) : 2.0f * s / (1.0f + r);   // axialStrain
                        sv = (1 + ax) * (posLink ? 1 : -1);
                    } else sv = posLink ? 1.0 : -1.0;
                    corner[cI][i] = (double)(float)(half * sv);
                }
            }
            appendf(t, "%.1f,%.1f,%.1f,%.1f,%.1f,%.1f,", corner[0][0] * vs, corner[0][1] * vs, corner[0][2] * vs,
                    corner[1][0] * vs, corner[1][1] * vs, corner[1][2] * vs);
            appendf(t, "%d,", sh.matid[v]);
            appendf(t, "%.1f,", sig.empty() ? 0.0 : sig[v]);   // localSignal
            appendf(t, ";");
        }
        appendf(t, "<<<>>>");
    }
    if (record_link) {
        appendf(t, "|[[[%d]]]", step_label);
        for (int l = 0; l < NL; l++) {
            if (lflags[l] & LF_GONE) continue;   // detached links are not drawn
            const double* a = &pos[3 * (size_t)sh.lpos[l]];
            const double* z = &pos[3 * (size_t)sh.lneg[l]];
            appendf(t, "%.4f,%.4f,%.4f,", a[0], a[1], a[2]);
            appendf(t, "%.4f,%.4f,%.4f,", z[0], z[1], z[2]);
            appendf(t, ";");
        }
        appendf(t, "[[[]]]");
    }
    appendf(t, "\n");
    return putText(t, buf, cap, len);
}

extern "C" vx3h_status vx3h_build_topology(const vx3h_sim_desc* desc, vx3h_topology_view* v) {
    if (!desc || !v) FAILV(VX3H_EINVAL, "null argument");
    HostModel H;
    if (!buildModel(desc, H, g_err)) return VX3H_EINVAL;
    v->n_voxels = H.nv;
    v->n_links = H.nl;
    for (int i = 0; i < H.nl; i++) {
        if (v->link_vneg) v->link_vneg[i] = H.lneg[i];
        if (v->link_vpos) v->link_vpos[i] = H.lpos[i];
        if (v->link_axis) v->link_axis[i] = H.laxis[i];
    }
    if (v->voxel_links) memcpy(v->voxel_links, H.voxelLinks.data(), H.voxelLinks.size() * sizeof(int));
    return VX3H_OK;
}
