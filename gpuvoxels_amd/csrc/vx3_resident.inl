// vx3_resident.inl -- persistent multi-step kernel: one workgroup advances one
// simulation for many steps with its working set on chip.
//
// Each thread owns up to LPT links and VPT voxels for the whole launch and keeps
// their mutable state in registers.  What other threads need crosses LDS:
//   pose   position + orientation + thermal factor (+ Poisson strain) per voxel,
//          written by the voxel owner, read by the link owners;
//   loads  the two local end loads of every link, written by the link owner and
//          gathered by the voxel owners in direction order X_POS..Z_NEG, exactly
//          as the stream path and CVX_Voxel::force() sum them.
// A step is: link pass | barrier | voxel pass + per-simulation control | barrier.
// Global memory is read once at launch start and written once at the end in the
// stream path's layout, so the two paths can hand simulations to each other.
//
// Included by the per-precision kernel units after vx3_kernels.inl (VX_NS, real).

namespace vx3 {
namespace VX_NS {

template <int BS, int VPT, int LPT>
struct ResidentShape {
    static constexpr int VMAX = BS * VPT;
    static constexpr int LMAX = BS * LPT;
    // LDS bytes: pose (7 real + tfac + 3 poisson floats per voxel), loads (12 real per link),
    // end strains (2 floats per link)
    static constexpr size_t bytes = (size_t)VMAX * (7 * sizeof(real) + 4 * sizeof(float)) +
                                    (size_t)LMAX * (12 * sizeof(real) + 2 * sizeof(float));
};

template <int BS, int VPT, int LPT>
__global__ void __launch_bounds__(BS) k_resident(DevBatch<real> b, const int* simList, int nSteps) {
    typedef ResidentShape<BS, VPT, LPT> Shape;
    constexpr int VMAX = Shape::VMAX, LMAX = Shape::LMAX;
    extern __shared__ __align__(16) unsigned char smem[];
    real* sPx = (real*)smem;
    real* sPy = sPx + VMAX; real* sPz = sPy + VMAX;
    real* sOw = sPz + VMAX; real* sOx = sOw + VMAX; real* sOy = sOx + VMAX; real* sOz = sOy + VMAX;
    real* sF = sOz + VMAX;                      // [12][LMAX]
    float* sTf = (float*)(sF + 12 * LMAX);
    float* sPsx = sTf + VMAX; float* sPsy = sPsx + VMAX; float* sPsz = sPsy + VMAX;
    float* sEs = sPsz + VMAX;                   // [2][LMAX]
    unsigned short* sVL = (unsigned short*)sF;  // [6][VMAX] link table, only during setup
    __shared__ int sRun, sDiv;
    __shared__ double sQ[VX3_DEV_MT_MAX + 1];
    __shared__ double sRed[4][BS / 64];

    const int s = simList[blockIdx.x];
    SimCtl& scg = b.sims[s];
    const int tid = threadIdx.x;
    if (scg.state != 0) return;
    const int v0 = scg.v0, nv = scg.nv, l0 = scg.l0, nl = scg.nl;
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
    const size_t VC = (size_t)b.Vcap;

    // ---- setup: voxel link table, link state and constants into registers
    for (int i = tid; i < 6 * VMAX; i += BS) sVL[i] = 0xFFFF;
    __syncthreads();
    int lNeg[LPT], lPos[LPT], lMeta[LPT], lMatN[LPT], lMatP[LPT];
    LinkState<real> ls[LPT];
#pragma unroll
    for (int k = 0; k < LPT; k++) {
        const int l = tid + k * BS;
        lNeg[k] = -1;
        if (l < nl) {
            const int g = l0 + l;
            lNeg[k] = b.lneg[g] - v0;
            lPos[k] = b.lpos[g] - v0;
            lMeta[k] = b.lmeta[g];
            lMatN[k] = b.vmat[v0 + lNeg[k]];
            lMatP[k] = b.vmat[v0 + lPos[k]];
            ls[k].pos2 = V3<real>(b.q2x[g], b.q2y[g], b.q2z[g]);
            ls[k].a1v = V3<real>(b.a1x[g], b.a1y[g], b.a1z[g]);
            ls[k].a2v = V3<real>(b.a2x[g], b.a2y[g], b.a2z[g]);
            ls[k].strain = b.strain[g];
            ls[k].maxStrain = b.maxStrain[g];
            ls[k].strainOffset = b.strainOffset[g];
            ls[k].flags = b.lflags[g];
            const int ax = lMeta[k] & 3;
            sVL[(2 * ax) * VMAX + lNeg[k]] = (unsigned short)l;
            sVL[(2 * ax + 1) * VMAX + lPos[k]] = (unsigned short)l;
            if (poisson) {
                sEs[l] = b.es[(2 * ax) * VC + v0 + lNeg[k]];
                sEs[LMAX + l] = b.es[(2 * ax + 1) * VC + v0 + lPos[k]];
            }
        }
    }
    __syncthreads();
    V3<real> vp[VPT], vm[VPT], va[VPT];
    Q4<real> vo[VPT];
    int vf[VPT], vmat[VPT], vext[VPT];
    double vph[VPT];
    unsigned short vl[VPT][6];
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int v = tid + k * BS;
        vmat[k] = -1;
        if (v < nv) {
            const int g = v0 + v;
            vp[k] = V3<real>(b.px[g], b.py[g], b.pz[g]);
            vm[k] = V3<real>(b.mx[g], b.my[g], b.mz[g]);
            va[k] = V3<real>(b.ax[g], b.ay[g], b.az[g]);
            vo[k] = Q4<real>(b.ow[g], b.ox[g], b.oy[g], b.oz[g]);
            vf[k] = b.vflags[g];
            vmat[k] = b.vmat[g];
            vext[k] = b.vext[g];
            vph[k] = b.phase[g];
#pragma unroll
            for (int d = 0; d < 6; d++) vl[k][d] = sVL[d * VMAX + v];
        }
    }
    __syncthreads();   // sVL aliases sF: done with it
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        const int v = tid + k * BS;
        if (v < nv) {
            const int g = v0 + v;
            sPx[v] = vp[k].x; sPy[v] = vp[k].y; sPz[v] = vp[k].z;
            sOw[v] = vo[k].w; sOx[v] = vo[k].x; sOy[v] = vo[k].y; sOz[v] = vo[k].z;
            sTf[v] = b.tfac[g];
            sPsx[v] = b.psx[g]; sPsy[v] = b.psy[g]; sPsz[v] = b.psz[g];
        }
    }
    double t = scg.t;
    int steps = scg.steps;
    int angleSamples = scg.angleSamples;
    double com[3] = {scg.com[0], scg.com[1], scg.com[2]};
    double comH0[3] = {scg.comH0[0], scg.comH0[1], scg.comH0[2]};
    double comH1[3] = {scg.comH1[0], scg.comH1[1], scg.comH1[2]};
    double recentAngle = scg.recentAngle;
    int state = 0;
    if (tid == 0) { sRun = 1; sDiv = 0; }
    __syncthreads();

    for (int it = 0; it < nSteps; it++) {
        // ---- link pass
#pragma unroll
        for (int k = 0; k < LPT; k++) {
            if (lNeg[k] < 0) continue;
            const int l = tid + k * BS;
            const int vn = lNeg[k], vq = lPos[k];
            const int axis = lMeta[k] & 3;
            const LinkMatDev& mat = b.lmatTab[lMeta[k] >> 2];
            const VoxMatDev& mn = b.vmatTab[lMatN[k]];
            const VoxMatDev& mp = b.vmatTab[lMatP[k]];
            if (mn.fixed && mp.fixed) continue;
            const real restLength = (real)0.5 * (nom * (real)sTf[vn] + nom * (real)sTf[vq]);
            float area, tss;
            if (mat.nu != 0.0f) {
                float an, sn, ap, sp;
                // CVX_Voxel::transverseArea / transverseStrainSum from the pose cache
                auto tr = [&](int v, bool nuZero, float& ar, float& su) {
                    if (nuZero) { ar = size * size; su = 0.0f; return; }
                    const float sx = sPsx[v], sy = sPsy[v], sz = sPsz[v];
                    const real s2 = (real)(size * size);
                    if (axis == 0) { ar = (float)(s2 * ((real)1 + (real)sy) * ((real)1 + (real)sz)); su = sy + sz; }
                    else if (axis == 1) { ar = (float)(s2 * ((real)1 + (real)sx) * ((real)1 + (real)sz)); su = sx + sz; }
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
            const float dmN = 2 * mn.sqrtMass * mn.zetaInternal / (mn.fixed ? 0.0f : dtf);
            const float dmP = 2 * mp.sqrtMass * mp.zetaInternal / (mp.fixed ? 0.0f : dtf);
            const V3<real> pN(sPx[vn], sPy[vn], sPz[vn]), pP(sPx[vq], sPy[vq], sPz[vq]);
            const Q4<real> oN(sOw[vn], sOx[vn], sOy[vn], sOz[vn]), oP(sOw[vq], sOx[vq], sOy[vq], sOz[vq]);
            V3<real> fN, mN, fP, mP;
            linkForces<real>(mat, axis, pN, oN, pP, oP, restLength, area, tss, mat.nu != 0.0f, dmN, dmP, ls[k], fN, mN, fP, mP);
            sF[0 * LMAX + l] = fN.x; sF[1 * LMAX + l] = fN.y; sF[2 * LMAX + l] = fN.z;
            sF[3 * LMAX + l] = mN.x; sF[4 * LMAX + l] = mN.y; sF[5 * LMAX + l] = mN.z;
            sF[6 * LMAX + l] = fP.x; sF[7 * LMAX + l] = fP.y; sF[8 * LMAX + l] = fP.z;
            sF[9 * LMAX + l] = mP.x; sF[10 * LMAX + l] = mP.y; sF[11 * LMAX + l] = mP.z;
            if (poisson) {
                const float r = mat.strainRatio;
                sEs[l] = 2.0f * ls[k].strain / (1.0f + r);
                sEs[LMAX + l] = 2.0f * ls[k].strain * r / (1.0f + r);
            }
            if (ls[k].strain > 100) sDiv = 1;
        }
        __syncthreads();
        const bool diverged = sDiv != 0;
        steps++;
        const bool sample = !diverged && cycleStep > 0 && steps % cycleStep == 0;
        double psx = 0, psy = 0, psz = 0, psm = 0;
        const double tn = t + dt;
        // ---- voxel pass
        if (!diverged) {
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                if (vmat[k] < 0) continue;
                const int v = tid + k * BS;
                const VoxMatDev& m = b.vmatTab[vmat[k]];
                if (!m.fixed) {
                    V3<real> f(0, 0, 0), mo(0, 0, 0);
#pragma unroll
                    for (int d = 0; d < 6; d++) {
                        const int l = vl[k][d];
                        if (l == 0xFFFF) continue;
                        const int e = (d & 1) ? 6 : 0;   // odd direction: this voxel is the positive end
                        f.add(V3<real>(sF[(e + 0) * LMAX + l], sF[(e + 1) * LMAX + l], sF[(e + 2) * LMAX + l]));
                        mo.add(V3<real>(sF[(e + 3) * LMAX + l], sF[(e + 4) * LMAX + l], sF[(e + 5) * LMAX + l]));
                    }
                    const ExtDev* ext = vext[k] >= 0 ? &b.extTab[vext[k]] : nullptr;
                    int ix = 0, iy = 0, iz = 0;
                    if (ext) { ix = ext->ix; iy = ext->iy; iz = ext->iz; }
                    voxelStep<real>(m, ext, dtf, nom, ix, iy, iz, sTf[v], f, mo, nullptr, 0, 0, nullptr, vp[k], vm[k], vo[k], va[k], vf[k]);
                    sPx[v] = vp[k].x; sPy[v] = vp[k].y; sPz[v] = vp[k].z;
                    sOw[v] = vo[k].w; sOx[v] = vo[k].x; sOy[v] = vo[k].y; sOz[v] = vo[k].z;
                }
                if (vary && !m.fixed && !(m.thermalOnAfter > tn)) {
                    double T = tempAmp * sin((double)(2 * 3.1415926f) * (tn / tempPeriod + vph[k]));
                    if (!expansion && T > 0) T = 0;
                    const float temp = (float)T;
                    sTf[v] = 1 + temp * m.alphaCTE;
                }
                if (poisson) {
                    // CVX_Voxel::strain(true) for the next step from this step's end strains
                    float st[3] = {0, 0, 0};
                    int n[3] = {0, 0, 0};
#pragma unroll
                    for (int d = 0; d < 6; d++) {
                        const int l = vl[k][d];
                        if (l == 0xFFFF) continue;
                        st[d >> 1] += sEs[((d & 1) ? LMAX : 0) + l];
                        n[d >> 1]++;
                    }
                    const ExtDev* ext = vext[k] >= 0 ? &b.extTab[vext[k]] : nullptr;
                    bool tension[3];
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        if (n[i] == 2) st[i] *= 0.5f;
                        tension[i] = (n[i] == 2) || (ext && (n[i] == 1 && ((ext->dofFixed & (1 << i)) || ext->force[i] != 0)));
                    }
                    if (!(tension[0] && tension[1] && tension[2])) {
                        float add = 0;
#pragma unroll
                        for (int i = 0; i < 3; i++) if (tension[i]) add += st[i];
                        const float pw = (float)pow((double)(1.0f + add), (double)(-m.nu));
                        const float value = pw - 1.0f;
#pragma unroll
                        for (int i = 0; i < 3; i++) if (!tension[i]) st[i] = value;
                    }
                    sPsx[v] = st[0]; sPsy[v] = st[1]; sPsz[v] = st[2];
                }
                if (sample && m.isMeasured) {
                    const double mass = m.mass;
                    psx += (double)vp[k].x * mass; psy += (double)vp[k].y * mass; psz += (double)vp[k].z * mass;
                    psm += mass;
                }
            }
        }
        if (sample) {
            // FP64 center of mass of the measured voxels (VX3_VoxelyzeKernel.cu:477-493)
            for (int o = 32; o > 0; o >>= 1) {
                psx += __shfl_xor(psx, o, 64); psy += __shfl_xor(psy, o, 64);
                psz += __shfl_xor(psz, o, 64); psm += __shfl_xor(psm, o, 64);
            }
            if ((tid & 63) == 0) { sRed[0][tid >> 6] = psx; sRed[1][tid >> 6] = psy; sRed[2][tid >> 6] = psz; sRed[3][tid >> 6] = psm; }
            __syncthreads();
            double tot[4] = {0, 0, 0, 0};
            for (int w = 0; w < BS / 64; w++) {
#pragma unroll
                for (int q = 0; q < 4; q++) tot[q] += sRed[q][w];
            }
            angleSamples++;
#pragma unroll
            for (int q = 0; q < 3; q++) { comH0[q] = comH1[q]; comH1[q] = com[q]; com[q] = tot[3] == 0 ? 0.0 : (1.0 / tot[3]) * tot[q]; }
            const double* A = comH0; const double* B = comH1; const double* C = com;
            const bool bc = B[0] == C[0] && B[1] == C[1] && B[2] == C[2];
            const bool ab = A[0] == B[0] && A[1] == B[1] && A[2] == B[2];
            if (bc || ab || angleSamples < 3) recentAngle = 0;
            else {
                double d1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, d2[3] = {C[0] - B[0], C[1] - B[1], C[2] - B[2]};
                double dot = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
                double l1 = sqrt((A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]));
                double l2 = sqrt((B[0] - C[0]) * (B[0] - C[0]) + (B[1] - C[1]) * (B[1] - C[1]) + (B[2] - C[2]) * (B[2] - C[2]));
                recentAngle = acos(dot / (l1 * l2));
            }
        }
        // ---- per-simulation control (uniform across the workgroup)
        if (diverged) { state = 2; break; }
        t = tn;
        bool stop = maxSteps > 0 && steps >= maxSteps;
        if (!stop && stopLen > 0 && tid == 0) sRun = mathTreeEval(b.tokens + stopOff, stopLen, com[0], com[1], com[2], 0, t, recentAngle, 0, 0, nv, sQ) > 0 ? 0 : 1;
        __syncthreads();
        if (stop || !sRun) { state = 1; break; }
    }

    // ---- write back in the stream layout
#pragma unroll
    for (int k = 0; k < LPT; k++) {
        if (lNeg[k] < 0) continue;
        const int l = tid + k * BS, g = l0 + l;
        b.q2x[g] = ls[k].pos2.x; b.q2y[g] = ls[k].pos2.y; b.q2z[g] = ls[k].pos2.z;
        b.a1x[g] = ls[k].a1v.x; b.a1y[g] = ls[k].a1v.y; b.a1z[g] = ls[k].a1v.z;
        b.a2x[g] = ls[k].a2v.x; b.a2y[g] = ls[k].a2v.y; b.a2z[g] = ls[k].a2v.z;
        b.strain[g] = ls[k].strain;
        b.maxStrain[g] = ls[k].maxStrain;
        b.strainOffset[g] = ls[k].strainOffset;
        b.lflags[g] = ls[k].flags;
        if (poisson) {
            const int ax = lMeta[k] & 3;
            b.es[(2 * ax) * VC + v0 + lNeg[k]] = sEs[l];
            b.es[(2 * ax + 1) * VC + v0 + lPos[k]] = sEs[LMAX + l];
        }
    }
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        if (vmat[k] < 0) continue;
        const int v = tid + k * BS, g = v0 + v;
        b.px[g] = vp[k].x; b.py[g] = vp[k].y; b.pz[g] = vp[k].z;
        b.mx[g] = vm[k].x; b.my[g] = vm[k].y; b.mz[g] = vm[k].z;
        b.ax[g] = va[k].x; b.ay[g] = va[k].y; b.az[g] = va[k].z;
        b.ow[g] = vo[k].w; b.ox[g] = vo[k].x; b.oy[g] = vo[k].y; b.oz[g] = vo[k].z;
        b.vflags[g] = vf[k];
        b.tfac[g] = sTf[v];
        b.psx[g] = sPsx[v]; b.psy[g] = sPsy[v]; b.psz[g] = sPsz[v];
    }
    if (tid == 0) {
        scg.t = t;
        scg.steps = steps;
        scg.angleSamples = angleSamples;
#pragma unroll
        for (int q = 0; q < 3; q++) { scg.com[q] = com[q]; scg.comH0[q] = comH0[q]; scg.comH1[q] = comH1[q]; }
        scg.recentAngle = recentAngle;
        scg.state = state;
    }
}

}  // namespace VX_NS
}  // namespace vx3
