/*
 * vx_mathtree.h -- MathTree evaluator shared by the oracle and the reference
 * driver (TEST INFRASTRUCTURE ONLY).  Restates the queue evaluator of
 * Voxelyze3/src/Utils/VX3_MathTree.h:50-192: tokens are the reversed BFS of the
 * XML expression tree; values[] is consumed from the front (pc) while results
 * are appended at the back (vc).
 */
#ifndef VX_MATHTREE_H
#define VX_MATHTREE_H
#include <cmath>
#include <math.h>
#include <vector>
#include "../include/vx3h.h"

namespace vxo {
enum { mtEND, mtCONST, mtE, mtPI, mtVAR, mtADD, mtSUB, mtMUL, mtDIV, mtPOW, mtSQRT, mtSIN, mtCOS, mtTAN, mtATAN, mtLOG, mtINT,
       mtABS, mtNOT, mtGREATERTHAN, mtLESSTHAN, mtAND, mtOR, mtNORMALCDF };

static double mathtree_eval(const vx3h_token* buff, int n, double x, double y, double z, double hit, double t, double angle,
                            double closeness, int numClosePairs, int num_voxel) {
    if (n <= 0 || buff == nullptr) return 0.0;   // an absent program evaluates to 0
    std::vector<double> values(n + 1, 0.0);
    int vc = 0, pc = 0;
    for (int i = 0; i < n; i++) {
        const vx3h_token& k = buff[i];
        switch (k.op) {
        case mtEND: return values[pc];
        case mtCONST: values[vc] = k.value; break;
        case mtE: values[vc] = 2.71828182845904523536; break;
        case mtPI: values[vc] = 3.14159265358979323846; break;
        case mtVAR:
            if (k.value < 0.5) values[vc] = x;
            else if (k.value < 1.5) values[vc] = y;
            else if (k.value < 2.5) values[vc] = z;
            else if (k.value < 3.5) values[vc] = hit;
            else if (k.value < 4.5) values[vc] = t;
            else if (k.value < 5.5) values[vc] = angle;
            else if (k.value < 6.5) values[vc] = closeness;
            else if (k.value < 7.5) values[vc] = numClosePairs;
            else if (k.value < 8.5) values[vc] = num_voxel;
            break;
        case mtSIN: values[vc] = sin(values[pc]); pc++; break;
        case mtCOS: values[vc] = cos(values[pc]); pc++; break;
        case mtTAN: values[vc] = tan(values[pc]); pc++; break;
        case mtATAN: values[vc] = atan(values[pc]); pc++; break;
        case mtLOG: values[vc] = log(values[pc]); pc++; break;
        case mtINT: values[vc] = rint(values[pc]); pc++; break;
        case mtNORMALCDF: values[vc] = 0.5 * erfc(-values[pc] / sqrt(2.0)); pc++; break;
        case mtADD: values[vc] = values[pc + 1] + values[pc]; pc += 2; break;
        case mtSUB: values[vc] = values[pc + 1] - values[pc]; pc += 2; break;
        case mtMUL: values[vc] = values[pc + 1] * values[pc]; pc += 2; break;
        case mtDIV: values[vc] = values[pc + 1] / values[pc]; pc += 2; break;
        case mtPOW: values[vc] = pow(values[pc + 1], values[pc]); pc += 2; break;
        case mtSQRT: values[vc] = sqrt(values[pc]); pc++; break;
        case mtABS: values[vc] = fabs(values[pc]); pc++; break;
        case mtNOT: values[vc] = !values[pc]; pc++; break;
        case mtGREATERTHAN: values[vc] = values[pc + 1] > values[pc]; pc += 2; break;
        case mtLESSTHAN: values[vc] = values[pc + 1] < values[pc]; pc += 2; break;
        case mtAND: values[vc] = values[pc + 1] && values[pc]; pc += 2; break;
        case mtOR: values[vc] = values[pc + 1] || values[pc]; pc += 2; break;
        default: return -1;
        }
        if (pc > vc) return -1;
        vc++;
    }
    return values[pc];
}

}  // namespace vxo
#endif
