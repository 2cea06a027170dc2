// Checks planTiles (gpuvoxels_amd/csrc/vx3h_tile.cpp) on one simulation's topology read
// from stdin: "nv nl vmax lax maxTiles", nv lines "i j k", nl lines "neg pos axis".
// Prints "ok <tiles> <owned> <halo>" when every invariant the tiled kernel relies on
// holds, else "bad <reason>".
#include <cstdio>
#include <set>
#include <vector>

#include "../gpuvoxels_amd/csrc/vx3h_tile.h"

using namespace vx3;

#define BAD(...) do { printf("bad "); printf(__VA_ARGS__); printf("\n"); return 0; } while (0)

int main() {
    int nv, nl, vmax, lax, maxTiles;
    if (scanf("%d %d %d %d %d", &nv, &nl, &vmax, &lax, &maxTiles) != 5) BAD("header");
    std::vector<int> ijk(3 * nv), lneg(nl), lpos(nl), lax3(nl);
    for (int v = 0; v < nv; v++) if (scanf("%d %d %d", &ijk[3 * v], &ijk[3 * v + 1], &ijk[3 * v + 2]) != 3) BAD("voxel");
    for (int l = 0; l < nl; l++) if (scanf("%d %d %d", &lneg[l], &lpos[l], &lax3[l]) != 3) BAD("link");
    TilePlan p;
    if (!planTiles(nv, ijk.data(), nl, lneg.data(), lpos.data(), lax3.data(), vmax, lax, maxTiles, &p)) {
        printf("none\n");
        return 0;
    }
    const int T = (int)p.tiles.size();
    std::vector<int> owner(nv, -1), writes(nl, 0);
    long long owned = 0, halo = 0;
    for (int t = 0; t < T; t++) {
        const TileDesc& d = p.tiles[t];
        if (d.tile != t || d.nTiles != T) BAD("tile numbering");
        if (d.nOwn < 1 || d.nOwn + d.nHalo > vmax) BAD("tile %d holds %d + %d voxels", t, d.nOwn, d.nHalo);
        for (int i = 0; i < d.nOwn; i++) {
            const int raw = p.vox[d.vOff + i], v = raw & 0x7fffffff;
            if (i) {   // voxels other tiles read first, each group ascending
                const int pr = p.vox[d.vOff + i - 1];
                if ((raw & 0x80000000) && !(pr & 0x80000000)) BAD("a published voxel after an unpublished one");
                if (((raw ^ pr) & 0x80000000) == 0 && v <= (pr & 0x7fffffff)) BAD("owned voxels not ascending");
            }
            if (owner[v] >= 0) BAD("voxel %d owned twice", v);
            owner[v] = t;
        }
        owned += d.nOwn;
        halo += d.nHalo;
    }
    if (owned != nv) BAD("owned %lld of %d voxels", owned, nv);
    for (int t = 0; t < T; t++) {
        const TileDesc& d = p.tiles[t];
        std::set<int> want, wantHalo;
        std::set<int> pub;
        for (int l = 0; l < nl; l++) {
            const bool a = owner[lneg[l]] == t, b = owner[lpos[l]] == t;
            if (!a && !b) continue;
            want.insert(l);
            if (!a) wantHalo.insert(lneg[l]);
            if (!b) wantHalo.insert(lpos[l]);
            if (a != b) pub.insert(a ? lneg[l] : lpos[l]);
        }
        if ((int)want.size() != d.nLinks) BAD("tile %d has %d links, %zu touch it", t, d.nLinks, want.size());
        if ((int)wantHalo.size() != d.nHalo) BAD("tile %d halo %d, expected %zu", t, d.nHalo, wantHalo.size());
        int per[3] = {0, 0, 0};
        for (int j = 0; j < d.nLinks; j++) {
            const int raw = p.link[d.lOff + j], l = raw & 0x7fffffff;
            if (!want.count(l)) BAD("tile %d computes link %d it does not touch", t, l);
            per[lax3[l]]++;
            const bool w = (raw & 0x80000000) != 0;
            if (w != (owner[lneg[l]] == t)) BAD("writer bit of link %d in tile %d", l, t);
            writes[l] += w;
            const int e = p.ends[d.lOff + j], sn = e & 0xFFFF, sp = e >> 16;
            if (sn >= d.nOwn + d.nHalo || sp >= d.nOwn + d.nHalo) BAD("slot out of range");
            if ((p.vox[d.vOff + sn] & 0x7fffffff) != lneg[l] || (p.vox[d.vOff + sp] & 0x7fffffff) != lpos[l])
                BAD("link %d ends map to the wrong slots in tile %d", l, t);
        }
        if (per[0] > lax || per[1] > lax || per[2] > lax) BAD("tile %d links per axis %d %d %d", t, per[0], per[1], per[2]);
        for (int i = 0; i < d.nHalo; i++) {
            const int v = p.vox[d.vOff + d.nOwn + i];
            if (v & 0x80000000) BAD("halo voxel carries the publish bit");
            if (!wantHalo.count(v)) BAD("halo voxel %d not a partner", v);
        }
        for (int i = 0; i < d.nOwn; i++) {
            const int raw = p.vox[d.vOff + i];
            if (((raw & 0x80000000) != 0) != (pub.count(raw & 0x7fffffff) != 0)) BAD("publish bit of voxel %d", raw & 0x7fffffff);
        }
    }
    for (int l = 0; l < nl; l++) if (writes[l] != 1) BAD("link %d written back by %d tiles", l, writes[l]);
    // exchange records: each published voxel has one, a tile's in slot order, every halo slot
    // reads its owner's record of that voxel
    std::vector<int> recOf(nv, -1);
    int next = 0;
    for (int t = 0; t < T; t++) {
        const TileDesc& d = p.tiles[t];
        for (int i = 0; i < d.nOwn; i++) {
            const int raw = p.vox[d.vOff + i], x = p.xidx[d.vOff + i];
            if (raw & 0x80000000) {
                if (x != next) BAD("record %d of voxel %d, expected %d", x, raw & 0x7fffffff, next);
                recOf[raw & 0x7fffffff] = next++;
            } else if (x != -1) BAD("unpublished voxel %d has record %d", raw, x);
        }
    }
    if (next != p.nPub) BAD("nPub %d, %d records", p.nPub, next);
    for (int t = 0; t < T; t++) {
        const TileDesc& d = p.tiles[t];
        for (int i = 0; i < d.nHalo; i++) {
            const int v = p.vox[d.vOff + d.nOwn + i], x = p.xidx[d.vOff + d.nOwn + i];
            if (x < 0 || x != recOf[v]) BAD("halo voxel %d of tile %d reads record %d, its owner writes %d", v, t, x, recOf[v]);
        }
    }
    printf("ok %d %lld %lld\n", T, owned, halo);
    return 0;
}
