/* vx_voxcad_stop.h -- VoxCAD's stop-condition family and equilibrium mode, the driver
 * logic of CVX_Sim (VoxCAD/src/VX1/VX_Sim.cpp:862-891 TimeStep, 1015-1055
 * StopConditionMet, 1057-1160 UpdateStats, 1833-1837 KineticEDecreasing), shared by the
 * oracle (vx_oracle.cpp) and the driver of the reference's compiled engine
 * (ref_driver.cpp), which supply the per-voxel and per-link energies.
 *
 * TEST INFRASTRUCTURE ONLY: part of the CPU oracle. */
#ifndef VX_VOXCAD_STOP_H
#define VX_VOXCAD_STOP_H

#include <cmath>
#include <vector>

struct VoxcadStop {
    static const int HIST_LEN = 500;   // HISTORY_SIZE (VoxCAD/src/VX1/VX_Sim.h)
    int type = 0;                          // StopCondition (old/VX_Enums.h:55-63)
    double value = 0;                      // StopConditionValue
    bool eq = false;                       // equilibrium mode
    std::vector<double> KinEHistory = std::vector<double>(HIST_LEN, -1.0);   // [0] newest, -1 unset
    std::vector<double> TotEHistory = std::vector<double>(HIST_LEN, -1.0);
    std::vector<double> MaxMoveHistory = std::vector<double>(HIST_LEN, -1.0);

    bool kineticEDecreasing() const {
        const std::vector<double>& k = KinEHistory;
        return k[0] + k[1] + k[2] < k[3] + k[4] + k[5] && !(k[0] == 0 || k[1] == 0 || k[2] == 0 || k[3] == 0 || k[4] == 0 || k[5] == 0);
    }
    static void push(std::vector<double>& h, double x) { h.pop_back(); h.insert(h.begin(), x); }

    /* After a successful step of size dt: the statistics the stop condition asks for
     * (velocities also whenever collisions are on), equilibrium-mode zeroing, then the
     * history push of UpdateStats.  halt() stops every voxel; stats(vel, kin, strainE,
     * maxVel2, totKE, totSE) returns the largest squared voxel speed and the total
     * kinetic and strain energies, each summed in double over float terms in list order. */
    template <class Halt, class Stats> void afterStep(bool collisions, double dt, Halt halt, Stats stats) {
        const bool kin = type == 4 || type == 5 || eq;
        const bool strainE = type == 4;
        const bool vel = type == 6 || collisions;
        if (!kin && !strainE && !vel) return;   // StatToCalc == CALCSTAT_NONE
        if (eq && kineticEDecreasing()) halt();  // ZeroAllMotion
        double maxVel2 = 0, totKE = 0, totSE = 0;
        stats(vel, kin, strainE, maxVel2, totKE, totSE);
        push(MaxMoveHistory, vel ? sqrt(maxVel2) * dt : -1.0);
        push(KinEHistory, kin ? totKE : -1.0);
        push(TotEHistory, (strainE && kin) ? totKE + totSE : -1.0);
    }

    /* StopConditionMet after `steps` steps at time t */
    bool met(long steps, double t, double tempPeriod, int nVoxels) const {
        if (steps < 2 * HIST_LEN) return false;
        const int numJump = HIST_LEN / 10;
        const double fNumVoxInv = 1.0 / (float)nVoxels;
        switch (type) {
        case 1: return steps > (int)(value + 0.5);
        case 2: return t > value;
        case 3: return t > tempPeriod * value;
        case 4: {
            const double ini = TotEHistory[0];
            for (int i = numJump; i < HIST_LEN; i += numJump) {
                if (TotEHistory[i] == -1) return false;
                if (fabs(TotEHistory[i] - ini) * fNumVoxInv > 0.001 * value) return false;
            }
            return true;
        }
        case 5:
            for (int i = 0; i < HIST_LEN; i += numJump) {
                if (KinEHistory[i] == -1) return false;
                if (KinEHistory[i] * fNumVoxInv > 0.001 * value) return false;
            }
            return true;
        case 6:
            for (int i = 0; i < HIST_LEN; i += numJump) {
                if (MaxMoveHistory[i] == -1) return false;
                if (MaxMoveHistory[i] > 0.001 * value) return false;
            }
            return true;
        default: return false;
        }
    }
};

#endif
