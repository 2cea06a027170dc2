// vx3h_tile.h -- cutting a simulation into tiles for the tiled resident kernel.
#ifndef VX3H_TILE_H
#define VX3H_TILE_H
#include <vector>
#include "vx3_types.h"

namespace vx3 {

// The tiles of one simulation.  Per tile (TileDesc, sim-local offsets into the arrays
// below; sim, xOff, cOff and pOff are the caller's to fill):
//   vox[vOff ..)   owned voxels (bit 31: a voxel of another tile links to it), those with
//                  bit 31 first, each group ascending; then the halo voxels ascending;
//   link[lOff ..)  every link with an owned end, ascending (bit 31: this tile owns the
//                  link's negative end and writes the link back);
//   ends[lOff ..)  the LDS slots of each such link's ends, negative | positive << 16;
//   xidx[vOff ..)  per slot its record in the simulation's exchange area: an owned voxel
//                  another tile reads writes record xidx (the tile's records are
//                  consecutive, in slot order), a halo voxel reads the record its
//                  owner writes; -1 for owned voxels no other tile reads.
// nPub: records in the simulation's exchange area (per step parity).
struct TilePlan {
    std::vector<TileDesc> tiles;
    std::vector<int> vox, link, ends, xidx;
    int nPub = 0;
};

// Cuts the voxels [0, nv) (lattice indices ijk, 3 per voxel) joined by links
// (lneg, lpos, laxis per link) into compact blocks by recursive coordinate bisection,
// each holding at most vmax owned + halo voxels and at most lax links per axis.
// Tries a few target block sizes and keeps the plan with the fewest tiles.  False when
// no plan has at most maxTiles tiles.
bool planTiles(int nv, const int* ijk, int nl, const int* lneg, const int* lpos, const int* laxis, int vmax, int lax,
               int maxTiles, TilePlan* out);

}  // namespace vx3
#endif
