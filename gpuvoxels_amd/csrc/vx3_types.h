// vx3_types.h -- POD tables shared by the host builder and the HIP kernels.
//
// Device layout (structure of arrays, one batch = many simulations):
//   voxels  [0, Vcap): position, linear momentum, orientation, angular momentum in
//           the state precision T (double: the reference's trajectories over the
//           full horizon; float: faster, same trajectories over short horizons),
//           plus flags, thermal size factor, material,
//           owning simulation.  Each simulation owns one contiguous voxel range.
//   links   [0, Lcap): endpoints, axis/material, beam state (previous local pose
//           for damping, strains, small-angle flags).  Contiguous range per sim.
//   forces  F[36][Vcap]: the link pass writes each link end's local force and
//           moment into the slot (direction, voxel) of the voxel it acts on; the
//           voxel pass reads its own six slots.  Slot d of voxel v is the link
//           in direction d (X_POS..Z_NEG, Voxelyze3/src/old/types.h:13-20) -- the
//           slot address is the gather index, so no CSR arrays or atomics are needed.
#ifndef VX3_TYPES_H
#define VX3_TYPES_H

#include <stdint.h>

#define VX3_MAX_CURVE 16

// voxel flag bits (state bits follow Voxelyze3/src/old/types.h:32-38)
enum {
    VF_INTERIOR = 1 << 1,            // all six links present (the reference's "SURFACE" bit)
    VF_FLOOR = 1 << 2,
    VF_STATIC = 1 << 3,              // floor static friction
    VF_REMOVED = 1 << 4,             // removed by a secondary experiment: frozen, no links
    VF_COLL = 1 << 5,
    VF_TARGET = 1 << 6,              // material isTarget: contacts with other materials count as hits
    VF_ATTACH = 1 << 7,              // VX3 enableAttach: the attach conditions held at the last update
    VF_LINKMASK_SHIFT = 8,           // bits 8..13: link present in direction d
};
// link state bits
enum { LF_LOCAL_VELOCITY_VALID = 1, LF_SMALL_ANGLE = 2,
       LF_GONE = 4,     // detached: no longer updated, no longer in its voxels' link masks
       LF_FRESH = 8 };  // made by attach this step: no force until the next step's link pass
// per-sim control bits
enum { SC_VARY = 1, SC_EXPANSION = 2, SC_DIVERGED_PENDING = 4, SC_POISSON = 8, SC_FLOOR = 16, SC_COLLISIONS = 32, SC_RESIDENT = 64,
       SC_REBUILD = 128,        // collision neighbour lists are rebuilt this step
       SC_COL_OVERFLOW = 256,   // a collision table overflowed: the simulation fails
       SC_EXTRAS = 512,         // cilia or a force field add to the voxel force (xf arrays)
       SC_CILIA = 1024,
       SC_DETACH = 2048,        // failed links detach
       SC_SECONDARY = 4096,     // secondary experiment: timed voxel removal, initial-position reset
       SC_SIGNALS = 8192,       // signals between linked voxels (runs with SC_EXTRAS)
       SC_TARGETS = 16384,      // target contacts are counted (runs with SC_EXTRAS)
       SC_ATTACH = 32768,       // touching sticky surface voxels get new links
       SC_COL_GROW = 65536,     // a collision table is too small: the simulation stands still
                                // (no further pass changes it) until the host enlarges the table
       SC_TOPOCOL = 1 << 17,    // collisions with topology changes (attach, detach, voxel removal):
                                // the nearby lists are regenerated on the device after a change
       SC_TOPO_CHG = 1 << 18,   // the topology changed during this step
       SC_NEARDIRTY = 1 << 19,  // this step regenerates the nearby lists from the current links
       SC_LINKS_DONE = 1 << 20,    // stood still for a collision table after this step's link pass:
                                   // the step resumes after the link pass (its end loads are kept)
       SC_NEARMASK = 1 << 21 };    // nearby sets also held as lattice-offset bit masks (nearMask)
#define VX3_COL_HBLOCK 1024     // hash buckets per scan block; a simulation's buckets are whole blocks

// A voxel's nearby set (the surface voxels within the link-hop radius) as a bit mask over
// lattice offsets: the voxels of one simulation sit at distinct lattice points and a voxel
// h hops away is at most h lattice steps away, so with a radius of at most 5 hops every
// member is at one of the 230 offsets (dx, dy, dz) with 1 <= |dx|+|dy|+|dz| <= 5.
#define VX3_NEAR_RADIUS 5
#define VX3_NEAR_WORDS 8        // 230 bits
struct NearBitTable {
    short bit[11 * 11 * 11];    // by (dx+5)*121 + (dy+5)*11 + (dz+5): the offset's bit, or -1
    constexpr NearBitTable() : bit() {
        int n = 0;
        for (int dx = -5; dx <= 5; dx++)
            for (int dy = -5; dy <= 5; dy++)
                for (int dz = -5; dz <= 5; dz++) {
                    const int l = (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy) + (dz < 0 ? -dz : dz);
                    bit[(dx + 5) * 121 + (dy + 5) * 11 + (dz + 5)] = (l == 0 || l > 5) ? (short)-1 : (short)n++;
                }
    }
};

struct VoxMatDev {                  // CVX_MaterialVoxel derived constants (VX_MaterialVoxel.cpp:57-79)
    float mass, massInverse, sqrtMass, momentInertiaInverse;
    float c2xSqMxExS, c2xSqIxExSxSxS;
    float zetaInternal, zetaGlobal, zetaCollision;
    float gravityForce;              // -mass*9.80665f*gravMult
    float penetrationStiffness;      // (float)(2*E*nomSize)
    float muStatic, muKinetic;
    float alphaCTE;
    float nu;
    float E;
    double nomSize;
    double thermalOnAfter;
    int fixed;                       // VX3 material-fixed
    int isMeasured;
    int sticky;                      // VX3 attach: 1 + the material's index in its simulation, 0 = not sticky
    int selfLink;                    // link-material table index of a link between two voxels of this material
};

struct LinkMatDev {                 // CVX_MaterialLink derived constants (VX_MaterialLink.cpp:121-141)
    float a1, a2, b1, b2, b3;
    float sqA1, sqA2xIp, sqB1, sqB2xFMp, sqB3xIp;
    float E, nu, eHat;
    float epsilonFail;               // -1 = none
    float strainRatio;               // E(pos voxel)/E(neg voxel)
    int linear;
    int nData;
    float strainData[VX3_MAX_CURVE];
    float stressData[VX3_MAX_CURVE];
};

struct ExtDev {                     // CVX_External (VX_External.h:34-94)
    int dofFixed;
    float force[3], moment[3];
    double translate[3];
    double rotQ[4];                  // rotationQuat() (w,x,y,z)
    int hasRotQ;
    int ix, iy, iz;
};

struct SimCtl {                     // one running simulation (VX3_VoxelyzeKernel scalars)
    double t;                        // currentTime (double accumulator)
    double optDt;                    // OptimalDt
    double tempAmp, tempPeriod;
    double dtFrac;
    double com[3], comH0[3], comH1[3], com0[3];
    double recentAngle;
    double voxSize;
    float dt;                        // step size (float, as doTimeStep's parameter)
    float prevDt;                    // the last step's dt: CVX_Voxel::previousDt of the link damping
                                     // (VX_Voxel.h:131); 0 before the first step
    float tInit;                     // temperature of voxels not driven by the actuation
    int state;                       // VX3H_SIM_*
    int steps, maxSteps, cycleStep, angleSamples;
    int flags;                       // SC_*
    int v0, nv, l0, nl;
    int stopOff, stopLen;            // MathTree program in the token pool
    int mat0, nvm, nlm;              // the sim's materials: simMats[mat0 .. +nvm) voxel, then nlm link table indices
    // collisions (CVoxelyze::updateCollisions, Voxelyze.cpp:671-711)
    float colThreshSq;               // watchRadiusMm^2 (pairs closer than this are watched)
    float colRecalcSq;               // recalcDist^2 (a surface voxel moving further triggers a rebuild)
    float colCell;                   // spatial-hash cell edge = watchRadiusMm
    int colInterval;                 // > 0: also rebuild every colInterval steps
    int colNearHops;                 // link-hop radius of the excluded "nearby" voxels (int(2*watchRadiusVx))
    int nb0, nbCap;                  // neighbour-list range in nbList
    int nbNeed;                      // entries the last rebuild needed when it set SC_COL_GROW
    int nbUsed, nearUsed;            // entries handed out so far by the current rebuild of each table
    int h0, hBits;                   // hash-table range: (1 << hBits) buckets from bucket h0 (whole blocks)
    int c0;                          // cell-sorted surface voxels: range [c0, c0 + nv) of cellVox
    int near0, nearCap, nearNeed;    // nearby-list range in nearList, and the size a rebuild needed
    int at0;                         // attach: attachable voxels and their pair offsets, attCand[at0 .. at0 + 2 nv)
    int ap0, apCap, apNeed;          // attach: candidate pair table attPair[ap0 .. ap0 + apCap), size last needed
    int rebuilds;                    // neighbour-list rebuilds so far
    // VoxCAD stop-condition family and equilibrium mode (VoxCAD/src/VX1/VX_Sim.cpp:862-891,1015-1160)
    int scType;                      // VX3H_SC_*
    int eqMode;                      // zero all motion whenever the kinetic energy falls
    double scValue;                  // StopConditionValue
    double dtFixed;                  // > 0: the fixed step (else DtFrac*OptimalDt) for the max-move history
    int histHead;                    // ring index of the newest entry of the 500-step histories
    // VX3 force field (VX3_Voxel.cu:178-184): MathTree programs per axis in the token pool
    int ffOff[3], ffLen[3];
    // VX3 secondary experiment: the initial-position reset (removal times are per voxel)
    double reinitAfter;
    int reinitDone;
    int hits;                        // collisionCount: target contacts so far (MathTree variable hit)
    // VX3 close pairs (computeTargetCloseness, VX3_VoxelyzeKernel.cu:545-562): radius (m, 0 = off),
    // the sim's target voxels tgtList[tg0 .. tg0 + nTg), and the values of the last sample
    double closeR, closeness;
    int closePairs, tg0, nTg;
    // VX3 attach: link range capacity (links are made at l0 + nl), new-link guard steps,
    // attach-condition programs in the token pool
    int nlMax, safety;
    int atOff[5], atLen[5];
    int scanDone;                    // bucket blocks of the current rebuild scanned so far (k_col_bscan)
};

// One tile of a simulation that the resident kernel runs on several workgroups (tiled
// launch): the tile owns a compact block of the simulation's voxels and computes every
// link with an end among them; the other ends (its halo) are read from the exchange
// area after each step's barrier.  A link whose ends two tiles own is computed by both
// from the same inputs, so both keep the same state; one of them writes it back.
struct TileDesc {
    int sim;                         // simulation slot
    int tile, nTiles;                // index of this tile in its simulation, and their number
    int vOff, nOwn, nHalo;           // tileVox[vOff ..): nOwn owned voxels then nHalo halo voxels
    int lOff, nLinks;                // tileLink / tileEnds[lOff .. lOff + nLinks)
    int xOff, xPar;                  // exchange area from xOff: two step parities of xPar reals, each
                                     // VX3_XCH_STRIDE reals per voxel another tile reads
    int cOff;                        // sync word cOff: the simulation's divergence word
    int pOff;                        // centre-of-mass partial sums: [2][nTiles][4] doubles from pOff
};
#define VX3_XCH_STRIDE 12            // position (3), orientation (4), thermal factor, Poisson strains (3), pad

#define VX3_HIST 500                 // HISTORY_SIZE (VoxCAD/src/VX1/VX_Sim.h)

struct TokenDev { int op; int pad; double value; };

#endif
