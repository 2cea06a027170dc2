// vx3h_tile.cpp -- cutting a simulation into tiles for the tiled resident kernel.
//
// A tile's workgroup holds the poses of its owned voxels and of their link partners in
// other tiles (the halo) in LDS, and computes every link with an owned end.  Compact
// blocks keep the halo small: recursive bisection across the longest lattice extent,
// splitting at the quantile that leaves each side a whole number of target blocks.
#include "vx3h_tile.h"

#include <algorithm>
#include <climits>
#include <cmath>

namespace vx3 {
namespace {

struct Planner {
    int nv = 0, nl = 0;
    const int *ijk = nullptr, *lneg = nullptr, *lpos = nullptr, *laxis = nullptr;
    int vmax = 0, lax = 0;
    std::vector<int> inc0, inc;   // links incident to each voxel (CSR)
    std::vector<int> vmark, lmark;
    int epoch = 0;
    bool failed = false;
    std::vector<std::vector<int>> parts;

    void index() {
        inc0.assign(nv + 1, 0);
        for (int l = 0; l < nl; l++) { inc0[lneg[l] + 1]++; inc0[lpos[l] + 1]++; }
        for (int v = 0; v < nv; v++) inc0[v + 1] += inc0[v];
        inc.resize(inc0[nv]);
        std::vector<int> at(inc0.begin(), inc0.end() - 1);
        for (int l = 0; l < nl; l++) { inc[at[lneg[l]]++] = l; inc[at[lpos[l]]++] = l; }
        vmark.assign(nv, -1);
        lmark.assign(nl, -1);
    }

    int other(int l, int v) const { return lneg[l] == v ? lpos[l] : lneg[l]; }

    // owned + halo voxels within vmax, links with an owned end within lax per axis
    int lastNeed = 0;   // slots (owned + halo) of the last block that did not fit
    bool fits(const int* part, size_t n) {
        lastNeed = (int)n;
        if ((int)n > vmax) return false;
        ++epoch;
        const int own = 2 * epoch, halo = own + 1;
        for (size_t i = 0; i < n; i++) vmark[part[i]] = own;
        int nh = 0, per[3] = {0, 0, 0};
        for (size_t j = 0; j < n; j++) {
            const int v = part[j];
            for (int i = inc0[v]; i < inc0[v + 1]; i++) {
                const int l = inc[i];
                if (lmark[l] == epoch) continue;
                lmark[l] = epoch;
                per[laxis[l]]++;
                const int u = other(l, v);
                if (vmark[u] != own && vmark[u] != halo) { vmark[u] = halo; nh++; }
            }
        }
        lastNeed = (int)n + nh;
        return (int)n + nh <= vmax && per[0] <= lax && per[1] <= lax && per[2] <= lax;
    }

    // k: the number of blocks the voxels part[0, n) are meant to become (the array is
    // reordered in place).  strict: a block that does not fit fails the plan (the caller
    // tries more blocks); else it is halved.
    bool strict = false;
    void cut(int* part, size_t n, int k) {
        if (failed || n == 0) return;
        if (k <= 1) {
            if (fits(part, n)) { parts.emplace_back(part, part + n); return; }
            if (strict || n == 1) { failed = true; return; }
            k = 2;
        }
        int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
        for (size_t i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], ijk[3 * part[i] + a]);
                hi[a] = std::max(hi[a], ijk[3 * part[i] + a]);
            }
        int ax = 0;
        for (int a = 1; a < 3; a++) if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
        const int kl = k / 2;
        size_t m = (size_t)((double)n * kl / k + 0.5);
        m = std::max<size_t>(1, std::min(m, n - 1));
        // the m lowest along the axis (ties by voxel index): a selection, not a sort
        std::nth_element(part, part + m, part + n, [&](int x, int y) {
            const int cx = ijk[3 * x + ax], cy = ijk[3 * y + ax];
            return cx != cy ? cx < cy : x < y;
        });
        cut(part, m, kl);
        cut(part + m, n - m, k - kl);
    }
};

}  // namespace

bool planTiles(int nv, const int* ijk, int nl, const int* lneg, const int* lpos, const int* laxis, int vmax, int lax,
               int maxTiles, TilePlan* out) {
    if (nv <= 0 || vmax <= 0 || lax <= 0 || maxTiles <= 0 || vmax > 0xFFFF) return false;
    Planner p;
    p.nv = nv; p.nl = nl; p.ijk = ijk; p.lneg = lneg; p.lpos = lpos; p.laxis = laxis; p.vmax = vmax; p.lax = lax;
    p.index();
    // Equal blocks: every tile's step takes as long as the largest block's, and each more
    // tile adds halo, exchange records and barrier arrivals, so the fewest blocks of equal
    // size that all fit.  From the count at which the blocks would fill every slot, a plan
    // that fails tries next the count at which its overflowing block's owned + halo would
    // have fit (blocks of n voxels need about n (1 + halo ratio) slots); the first count
    // that fits is then lowered one at a time towards the last that failed.  A shape no
    // equal split fits falls back to halving the blocks that do not.
    std::vector<std::vector<int>> best;
    std::vector<int> all(nv);
    auto attempt = [&](int k, bool strict) {
        for (int v = 0; v < nv; v++) all[v] = v;
        p.parts.clear();
        p.failed = false;
        p.strict = strict;
        p.cut(all.data(), all.size(), k);
        return !p.failed;
    };
    int failedK = 0;
    for (int k = std::max(1, (nv + vmax - 1) / vmax); k <= maxTiles;) {
        if (attempt(k, true)) {
            best = p.parts;
            for (int j = k - 1; j > failedK && attempt(j, true); j--) best = p.parts;
            break;
        }
        failedK = k;
        k = std::max(k + 1, (int)std::ceil((double)k * p.lastNeed / vmax));
    }
    if (best.empty())
        for (double frac : {0.8, 0.6}) {
            if (!attempt(std::max(1, (int)((nv + frac * vmax - 1) / (frac * vmax))), false)) continue;
            if (best.empty() || p.parts.size() < best.size()) best = p.parts;
        }
    if (best.empty() || (int)best.size() > maxTiles) return false;

    const int T = (int)best.size();
    std::vector<int> owner(nv, -1);
    for (int t = 0; t < T; t++)
        for (int v : best[t]) owner[v] = t;
    // per voxel: does a voxel of another tile link to it
    std::vector<char> shared(nv, 0);
    for (int l = 0; l < nl; l++)
        if (owner[lneg[l]] != owner[lpos[l]]) shared[lneg[l]] = shared[lpos[l]] = 1;
    // the links of each tile (those with an owned end), bucketed in link order
    std::vector<int> lcount(T + 1, 0);
    for (int l = 0; l < nl; l++) {
        lcount[owner[lneg[l]] + 1]++;
        if (owner[lpos[l]] != owner[lneg[l]]) lcount[owner[lpos[l]] + 1]++;
    }
    for (int t = 0; t < T; t++) lcount[t + 1] += lcount[t];
    std::vector<int> tlinks(lcount[T]);
    {
        std::vector<int> at(lcount.begin(), lcount.end() - 1);
        for (int l = 0; l < nl; l++) {
            tlinks[at[owner[lneg[l]]]++] = l;
            if (owner[lpos[l]] != owner[lneg[l]]) tlinks[at[owner[lpos[l]]]++] = l;
        }
    }
    out->tiles.clear(); out->vox.clear(); out->link.clear(); out->ends.clear(); out->xidx.clear();
    out->vox.reserve(nv + nv / 2); out->xidx.reserve(nv + nv / 2);
    out->link.reserve(tlinks.size()); out->ends.reserve(tlinks.size());
    std::vector<int> slot(nv, -1);
    std::vector<int> rec(nv, -1);   // exchange record of each voxel another tile reads
    int nPub = 0;
    std::vector<int> own, halo;
    for (int t = 0; t < T; t++) {
        std::vector<int>& part = best[t];
        std::sort(part.begin(), part.end());
        // owned voxels another tile reads take the lowest slots (each thread's first voxels:
        // their records are written first and drain while the rest integrate)
        own.clear();
        for (int v : part) if (shared[v]) own.push_back(v);
        for (int v : part) if (!shared[v]) own.push_back(v);
        halo.clear();
        for (int i = lcount[t]; i < lcount[t + 1]; i++) {
            const int l = tlinks[i];
            if (owner[lneg[l]] != t) halo.push_back(lneg[l]);
            if (owner[lpos[l]] != t) halo.push_back(lpos[l]);
        }
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        TileDesc d{};
        d.tile = t;
        d.nTiles = T;
        d.vOff = (int)out->vox.size();
        d.nOwn = (int)own.size();
        d.nHalo = (int)halo.size();
        d.lOff = (int)out->link.size();
        d.nLinks = lcount[t + 1] - lcount[t];
        for (int i = 0; i < d.nOwn; i++) {
            const int v = own[i];
            out->vox.push_back(v | (shared[v] ? (int)0x80000000u : 0));
            if (shared[v]) rec[v] = nPub++;
            out->xidx.push_back(shared[v] ? rec[v] : -1);
            slot[v] = i;
        }
        for (int i = 0; i < d.nHalo; i++) {
            out->vox.push_back(halo[i]);
            out->xidx.push_back(-2 - halo[i]);   // the owner's record, filled in below
            slot[halo[i]] = d.nOwn + i;
        }
        for (int i = lcount[t]; i < lcount[t + 1]; i++) {
            const int l = tlinks[i];
            out->link.push_back(l | (owner[lneg[l]] == t ? (int)0x80000000u : 0));
            out->ends.push_back(slot[lneg[l]] | (slot[lpos[l]] << 16));
        }
        for (int v : own) slot[v] = -1;
        for (int v : halo) slot[v] = -1;
        out->tiles.push_back(d);
    }
    for (int& x : out->xidx)
        if (x <= -2) x = rec[-2 - x];   // every halo voxel is linked to its reader: its owner publishes it
    out->nPub = nPub;
    return true;
}

}  // namespace vx3
