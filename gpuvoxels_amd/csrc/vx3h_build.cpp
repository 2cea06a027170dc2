// vx3h_build.cpp -- host model builder (see vx3h_build.h).
//
// Material derivation restates Voxelyze's setters and derived quantities:
//   CVX_Material setters / models     Voxelyze3/src/old/VX_Material.cpp:182-549
//   voxel material derived values     Voxelyze3/src/old/VX_MaterialVoxel.cpp:57-79
//   link material combination         Voxelyze3/src/old/VX_MaterialLink.cpp:45-141
//   importer call order               Voxelyze3/src/VXA/VX_Sim.cpp:59-153,368-420,445-483
// Topology restates CVoxelyze::addVoxel/addLink (Voxelyze3/src/old/Voxelyze.cpp:439-539)
// and CVX_Link's constructor (Voxelyze3/src/old/VX_Link.cpp:21-54).
//
// Float math: <math.h> is included so that sqrt/pow of a float argument are the
// single-precision overloads, as in the reference translation units.
#include "vx3h_build.h"

#include <cfloat>
#include <cmath>
#include <math.h>
#include <cstring>
#include <unordered_map>

namespace vx3 {

namespace {

struct Mat {
    bool linear = true;
    float E = 1.0f, sigmaYield = -1, sigmaFail = -1, epsYield = -1, epsFail = -1;
    std::vector<float> strain{0.f, 1e6f}, stress{0.f, 1e6f};
    float nu = 0, rho = 1, cte = 0, muS = 0, muK = 0, zi = 1, zg = 0, zc = 0, eHat = 1;
    double nomSize = 0.001;
    float gravMult = 0;
    bool fixed = false;
    int measured = 1;
    double thermalOn = 0;

    void derive() { eHat = E / ((1 - 2 * nu) * (1 + nu)); }
    bool failed(float s) const { return epsFail != -1.0f && s > epsFail; }
    float modulus(float s) const {
        if (failed(s)) return 0.0f;
        if (s <= strain[1] || linear) return E;
        const int n = (int)strain.size();
        for (int i = 2; i < n; i++)
            if (s <= strain[i] || i == n - 1) return (stress[i] - stress[i - 1]) / (strain[i] - strain[i - 1]);
        return 0.0f;
    }
    float strainAt(float st) const {
        if (st <= stress[1] || linear) return st / E;
        const int n = (int)strain.size();
        for (int i = 2; i < n; i++)
            if (st <= stress[i] || i == n - 1) {
                float p = (st - stress[i - 1]) / (stress[i] - stress[i - 1]);
                return strain[i - 1] + p * (strain[i] - strain[i - 1]);
            }
        return 0.0f;
    }
    bool modelLinear(float e, float fail = -1) {
        if (e <= 0) return false;
        if (fail != -1.0f && fail <= 0) return false;
        float fs = fail == -1 ? 1000000.0f : fail;
        float fe = fs / e;
        strain = {0.0f, fe};
        stress = {0.0f, fs};
        linear = true;
        E = e;
        sigmaYield = sigmaFail = fail;
        epsYield = epsFail = (fail == -1) ? -1.0f : fe;
        derive();
        return true;
    }
    bool modelBilinear(float e, float pm, float ys, float fail = -1) {
        if (e <= 0 || pm <= 0 || pm >= e || ys <= 0) return false;
        if (fail != -1.0f && fail <= ys) return false;
        float yStrain = ys / e;
        float fs = fail == -1 ? 3 * ys : fail;
        float b = ys - pm * yStrain;
        float fStrain = (fs - b) / pm;
        strain = {0.0f, yStrain, fStrain};
        stress = {0.0f, ys, fs};
        linear = false;
        E = e;
        sigmaYield = ys;
        sigmaFail = fail;
        epsYield = yStrain;
        epsFail = fail == -1.0f ? -1.0f : fStrain;
        derive();
        return true;
    }
    void yieldFromData(float pct = 0.2) {
        sigmaYield = epsYield = -1.0f;
        float oM = E, oB = (-pct / 100 * oM);
        int pts = (int)strain.size() - 1;
        for (int i = 1; i < pts - 1; i++) {
            float x1 = strain[i], x2 = strain[i + 1], y1 = stress[i], y2 = stress[i + 1];
            float tM = (y2 - y1) / (x2 - x1), tB = y1 - tM * x1;
            if (oM != tM) {
                float xi = (tB - oB) / (oM - tM);
                if (xi > x1 && xi < x2) {
                    sigmaYield = y1 + (xi - x1) / (x2 - x1) * (y2 - y1);
                    epsYield = xi;
                    return;
                }
            }
        }
        sigmaYield = sigmaFail;
        epsYield = epsFail;
    }
    bool model(int n, const float* st, const float* ss) {
        if (*st == 0 && *ss == 0) { st++; ss++; n--; }
        if (n <= 0 || *st <= 0 || *ss <= 0) return false;
        std::vector<float> a{0.0f}, b{0.0f};
        float sw = 0, ssw = 0;
        for (int i = 0; i < n; i++) {
            if (st[i] <= sw) return false;
            if (i > 0 && (ss[i] - ssw) / (st[i] - sw) > b[0] / a[0]) return false;
            sw = st[i];
            ssw = ss[i];
            a.push_back(st[i]);
            b.push_back(ss[i]);
        }
        strain = a;
        stress = b;
        E = stress[1] / strain[1];
        sigmaFail = stress.back();
        epsFail = strain.back();
        linear = (n == 1);
        if (n == 1 || n == 2) { sigmaYield = stress[1]; epsYield = strain[1]; }
        else yieldFromData();
        derive();
        return true;
    }
    void setNu(float p) { if (p < 0) p = 0; if (p >= 0.5) p = 0.5 - FLT_EPSILON * 2; nu = p; derive(); }
};

// CVX_MaterialVoxel::updateDerived
void voxelDerived(const Mat& m, VoxMatDev& o) {
    memset(&o, 0, sizeof(o));
    double vol = m.nomSize * m.nomSize * m.nomSize;
    o.mass = (float)(vol * m.rho);
    float inertia = (float)(o.mass * m.nomSize * m.nomSize / 6.0f);
    if (vol != 0 && o.mass != 0 && inertia != 0) {
        o.massInverse = 1.0f / o.mass;
        o.sqrtMass = sqrt(o.mass);
        o.momentInertiaInverse = 1.0f / inertia;
        o.c2xSqMxExS = (float)(2.0f * sqrt(o.mass * m.E * m.nomSize));
        o.c2xSqIxExSxSxS = (float)(2.0f * sqrt(inertia * m.E * m.nomSize * m.nomSize * m.nomSize));
    }
    o.zetaInternal = m.zi;
    o.zetaGlobal = m.zg;
    o.zetaCollision = m.zc;
    o.gravityForce = -o.mass * 9.80665f * m.gravMult;
    o.penetrationStiffness = (float)(2 * m.E * m.nomSize);
    o.muStatic = m.muS;
    o.muKinetic = m.muK;
    o.alphaCTE = m.cte;
    o.nu = m.nu;
    o.E = m.E;
    o.nomSize = m.nomSize;
    o.thermalOnAfter = m.thermalOn;
    o.fixed = m.fixed ? 1 : 0;
    o.isMeasured = m.measured;
}

// CVX_MaterialLink::updateAll + updateDerived, vox1 = material of the voxel that created the link.
Mat combine(const Mat& v1, const Mat& v2) {
    Mat l;   // CVX_MaterialVoxel() defaults: linear E = 1e6, rho = 1e3
    l.modelLinear(1e6f);
    l.rho = 1e3f;
    l.nomSize = 0.5 * (v1.nomSize + v2.nomSize);
    l.rho = 0.5f * (v1.rho + v2.rho);
    l.cte = 0.5f * (v1.cte + v2.cte);
    l.muS = 0.5f * (v1.muS + v2.muS);
    l.muK = 0.5f * (v1.muK + v2.muK);
    l.zi = 0.5f * (v1.zi + v2.zi);
    l.zg = 0.5f * (v1.zg + v2.zg);
    l.zc = 0.5f * (v1.zc + v2.zc);
    float f1 = v1.sigmaFail, f2 = v2.sigmaFail, fail;
    if (f1 == -1.0f) fail = f2;
    else if (f2 == -1.0f) fail = f1;
    else fail = f1 < f2 ? f1 : f2;
    if (v1.linear && v2.linear) {
        l.modelLinear(2.0f * v1.E * v2.E / (v1.E + v2.E), fail);
    } else {
        std::vector<float> ss{0.0f}, st{0.0f};
        int i1 = 1, i2 = 1;
        while (i1 < (int)v1.strain.size() && i2 < (int)v2.strain.size()) {
            float s = FLT_MAX;
            if (i1 < (int)v1.strain.size()) s = v1.strain[i1];
            if (i2 < (int)v2.strain.size() && v2.strain[i2] < s) s = v2.strain[i2];
            if (s == v1.strain[i1]) i1++;
            if (s == v2.strain[i2]) i2++;
            float m1 = v1.modulus(s - FLT_EPSILON), m2 = v2.modulus(s - FLT_EPSILON);
            float mm = 2.0f * m1 * m2 / (m1 + m2);
            int last = (int)st.size() - 1;
            st.push_back(s);
            ss.push_back(ss[last] + mm * (s - st[last]));
        }
        l.model((int)st.size(), st.data(), ss.data());
        l.sigmaFail = fail;
        l.epsFail = fail == -1.0f ? -1.0f : l.strainAt(fail);
    }
    if (v1.nu == 0 && v2.nu == 0) l.nu = 0;
    else {
        float eh = 2 * v1.eHat * v2.eHat / (v1.eHat + v2.eHat);
        float e = l.E;
        float c2 = (eh - e) / (2 * eh) + 0.0625;
        l.nu = sqrt(c2) - 0.25;
    }
    l.derive();
    return l;
}

void linkDerived(const Mat& l, LinkMatDev& o) {
    memset(&o, 0, sizeof(o));
    float L = (float)l.nomSize;
    o.a1 = l.E * L;
    o.a2 = l.E * L * L * L / (12.0f * (1 + l.nu));
    o.b1 = l.E * L;
    o.b2 = l.E * L * L / 2.0f;
    o.b3 = l.E * L * L * L / 6.0f;
    o.sqA1 = sqrt(o.a1);
    o.sqA2xIp = sqrt(o.a2 * L * L / 6.0f);
    o.sqB1 = sqrt(o.b1);
    o.sqB2xFMp = sqrt(o.b2 * L / 2.0f);
    o.sqB3xIp = sqrt(o.b3 * L * L / 6.0f);
    o.E = l.E;
    o.nu = l.nu;
    o.eHat = l.eHat;
    o.epsilonFail = l.epsFail;
    o.linear = l.linear ? 1 : 0;
    o.nData = (int)l.strain.size();
    for (int i = 0; i < o.nData && i < VX3_MAX_CURVE; i++) {
        o.strainData[i] = l.strain[i];
        o.stressData[i] = l.stress[i];
    }
}

struct KeyHash {
    size_t operator()(const long long& k) const { return std::hash<long long>()(k); }
};
long long key3(int x, int y, int z) {
    return ((long long)(x + (1 << 20)) << 42) | ((long long)(y + (1 << 20)) << 21) | (long long)(z + (1 << 20));
}

}  // namespace

bool buildModel(const vx3h_sim_desc* d, HostModel& H, std::string& err) {
    if (!d) { err = "null description"; return false; }
    if (d->n_materials <= 0) { err = "no materials"; return false; }
    if (d->n_voxels < 0 || (d->n_voxels > 0 && (!d->ijk || !d->material))) { err = "bad voxel arrays"; return false; }
    if (!(d->voxel_size > 0)) { err = "voxel_size must be positive"; return false; }
    const int NM = d->n_materials;
    std::vector<Mat> mats(NM);
    // CVX_Sim::Import: addMaterial() -> CopyMat -> damping setters (VX_Sim.cpp:77-86, 368-420)
    for (int i = 0; i < NM; i++) {
        const vx3h_material_desc& s = d->materials[i];
        Mat& m = mats[i];
        m.modelLinear(1e6f);   // CVoxelyze::addMaterial(1e6, 1e3)
        m.rho = 1e3f;
        m.derive();
        switch (s.mat_model) {
        case VX3H_MDL_LINEAR: m.modelLinear((float)s.elastic_mod); break;
        case VX3H_MDL_LINEAR_FAIL: m.modelLinear((float)s.elastic_mod, (float)s.fail_stress); break;
        case VX3H_MDL_BILINEAR:
            m.modelBilinear((float)s.elastic_mod, (float)s.plastic_mod, (float)s.yield_stress, (float)s.fail_stress);
            break;
        case VX3H_MDL_DATA: {
            if (s.n_data <= 0 || s.n_data > VX3H_MAX_CURVE - 1) { err = "MDL_DATA point count out of range"; return false; }
            std::vector<float> st(s.n_data), ss(s.n_data);
            for (int k = 0; k < s.n_data; k++) { st[k] = (float)s.strain_data[k]; ss[k] = (float)s.stress_data[k]; }
            m.model(s.n_data, st.data(), ss.data());
            break;
        }
        default: err = "unknown material model"; return false;
        }
        m.setNu((float)s.poissons_ratio);
        float rho = (float)s.density;
        if (rho <= 0) rho = FLT_MIN;
        m.rho = rho;
        m.cte = (float)s.cte;
        m.muS = (float)s.u_static > 0 ? (float)s.u_static : 0.0f;
        m.muK = (float)s.u_dynamic > 0 ? (float)s.u_dynamic : 0.0f;
        m.zi = (float)s.zeta_internal > 0 ? (float)s.zeta_internal : 0.0f;
        m.zg = (float)s.zeta_global > 0 ? (float)s.zeta_global : 0.0f;
        m.zc = (float)s.zeta_collision > 0 ? (float)s.zeta_collision : 0.0f;
        m.gravMult = d->gravity_mult;
        m.fixed = s.fixed != 0;
        m.measured = s.is_measured;
        m.thermalOn = s.thermal_on_after;
        m.nomSize = d->voxel_size;   // CVoxelyze::setVoxelSize
        // CVX_Sim::EnableVolumeEffects (VX_Sim.cpp:445-458)
        m.setNu(d->volume_effects ? (float)s.poissons_ratio : 0.0f);
        m.derive();
    }
    H = HostModel();
    H.voxSize = d->voxel_size;
    H.vmats.resize(NM);
    for (int i = 0; i < NM; i++) voxelDerived(mats[i], H.vmats[i]);

    // voxels and links in insertion order
    const int NV = d->n_voxels;
    H.nv = NV;
    H.ijk.assign(d->ijk, d->ijk + 3 * NV);
    H.vmat.assign(d->material, d->material + NV);
    H.phase.resize(NV);
    H.voxelLinks.assign(6 * NV, -1);
    std::unordered_map<long long, int, KeyHash> where;
    where.reserve(NV * 2 + 1);
    std::vector<std::pair<int, int>> pairMats;   // (vox1Mat, vox2Mat) per link material
    std::vector<Mat> linkMatsHost;
    static const int dx[6] = {1, -1, 0, 0, 0, 0}, dy[6] = {0, 0, 1, -1, 0, 0}, dz[6] = {0, 0, 0, 0, 1, -1};
    for (int v = 0; v < NV; v++) {
        int m = d->material[v];
        if (m < 0 || m >= NM) { err = "voxel material index out of range"; return false; }
        int x = d->ijk[3 * v], y = d->ijk[3 * v + 1], z = d->ijk[3 * v + 2];
        long long k = key3(x, y, z);
        if (where.count(k)) { err = "duplicate voxel position"; return false; }
        where[k] = v;
        H.phase[v] = d->phase ? d->phase[v] : 0.0;
        for (int dir = 0; dir < 6; dir++) {
            auto it = where.find(key3(x + dx[dir], y + dy[dir], z + dz[dir]));
            if (it == where.end()) continue;
            int u = it->second;
            // combinedMaterial(voxel1 mat = new voxel, voxel2 mat = neighbour) (Voxelyze.cpp:626-641)
            int m1 = m, m2 = d->material[u], lm = -1;
            for (int j = 0; j < (int)pairMats.size(); j++)
                if ((pairMats[j].first == m1 && pairMats[j].second == m2) || (pairMats[j].first == m2 && pairMats[j].second == m1)) {
                    lm = j;
                    break;
                }
            if (lm < 0) {
                lm = (int)pairMats.size();
                pairMats.push_back({m1, m2});
                linkMatsHost.push_back(combine(mats[m1], mats[m2]));
            }
            // CVX_Link ctor: the lower lattice coordinate is the negative end
            bool newIsNeg = (dir % 2) == 0;   // neighbour in the positive direction
            int vneg = newIsNeg ? v : u, vpos = newIsNeg ? u : v;
            int li = (int)H.lneg.size();
            H.lneg.push_back(vneg);
            H.lpos.push_back(vpos);
            H.laxis.push_back(dir / 2);
            H.lmat.push_back(lm);
            H.voxelLinks[6 * v + dir] = li;
            H.voxelLinks[6 * u + (dir ^ 1)] = li;
        }
    }
    H.nl = (int)H.lneg.size();
    // device link materials: one entry per (link material, which voxel material sits at the negative end)
    std::vector<int> variant;   // key lm*NM + negMat -> index
    std::unordered_map<int, int> vidx;
    for (int l = 0; l < H.nl; l++) {
        int negMat = H.vmat[H.lneg[l]], posMat = H.vmat[H.lpos[l]];
        int key = H.lmat[l] * NM + negMat;
        auto it = vidx.find(key);
        if (it == vidx.end()) {
            LinkMatDev o;
            linkDerived(linkMatsHost[H.lmat[l]], o);
            o.strainRatio = mats[posMat].E / mats[negMat].E;   // CVX_Link::reset (VX_Link.cpp:62)
            vidx[key] = (int)H.lmats.size();
            H.lmats.push_back(o);
            it = vidx.find(key);
        }
        H.lmat[l] = it->second;
    }
    // VX3 attach makes links between two voxels of one sticky material at run time
    // (VX3_Link device ctor, VX3_Link.cu:30-55: combinedMaterial(m, m), strain ratio 1)
    H.vmSelf.assign(NM, -1);
    if (d->enable_attach)
        for (int m = 0; m < NM; m++) {
            if (!d->materials[m].sticky || d->materials[m].fixed) continue;
            int lm = -1;
            for (int j = 0; j < (int)pairMats.size(); j++)
                if (pairMats[j].first == m && pairMats[j].second == m) { lm = j; break; }
            if (lm < 0) {
                lm = (int)pairMats.size();
                pairMats.push_back({m, m});
                linkMatsHost.push_back(combine(mats[m], mats[m]));
            }
            const int key = lm * NM + m;
            auto it = vidx.find(key);
            if (it == vidx.end()) {
                LinkMatDev o;
                linkDerived(linkMatsHost[lm], o);
                o.strainRatio = mats[m].E / mats[m].E;
                vidx[key] = (int)H.lmats.size();
                H.lmats.push_back(o);
                it = vidx.find(key);
            }
            H.vmSelf[m] = it->second;
            H.vmats[m].sticky = m + 1;
        }

    // flags: floor, static friction set by CVX_Voxel::reset, interior, link mask
    H.vflags.assign(NV, 0);
    for (int v = 0; v < NV; v++) {
        int f = VF_STATIC;
        if (d->floor_enabled) f |= VF_FLOOR;
        if (d->collisions_enabled) f |= VF_COLL;
        bool interior = true;
        for (int dir = 0; dir < 6; dir++) {
            if (H.voxelLinks[6 * v + dir] >= 0) f |= 1 << (VF_LINKMASK_SHIFT + dir);
            else interior = false;
        }
        if (interior) f |= VF_INTERIOR;
        H.vflags[v] = f;
    }

    // externals (boundary conditions resolved per voxel; VX_Sim.cpp:109-143, VX_External.cpp:45-112)
    H.vext.assign(NV, -1);
    for (int e = 0; e < d->n_externals; e++) {
        const vx3h_external& x = d->externals[e];
        if (x.voxel < 0 || x.voxel >= NV) { err = "external voxel index out of range"; return false; }
        int idx = H.vext[x.voxel];
        if (idx < 0) {
            ExtDev z;
            memset(&z, 0, sizeof(z));
            z.rotQ[0] = 1;
            z.ix = d->ijk[3 * x.voxel];
            z.iy = d->ijk[3 * x.voxel + 1];
            z.iz = d->ijk[3 * x.voxel + 2];
            idx = (int)H.exts.size();
            H.exts.push_back(z);
            H.vext[x.voxel] = idx;
        }
        ExtDev& E = H.exts[idx];
        double rot[3] = {0, 0, 0};
        for (int k = 0; k < 6; k++) {
            int bit = 1 << k;
            if (!(x.dof_fixed & bit)) continue;
            E.dofFixed |= bit;
            double val = k < 3 ? x.translate[k] : x.rotate[k - 3];
            if (val != 0.0f) {
                if (k < 3) E.translate[k] = val;
                else rot[k - 3] = val;
            }
        }
        for (int k = 0; k < 3; k++) {
            E.force[k] += (float)x.force[k];
            E.moment[k] += (float)x.moment[k];
        }
        if (rot[0] != 0 || rot[1] != 0 || rot[2] != 0) {
            // Quat3D(rotation vector): FromRotationVector (Quat3D.h:128-143)
            double th[3] = {0.5 * rot[0], 0.5 * rot[1], 0.5 * rot[2]};
            double m2 = th[0] * th[0] + th[1] * th[1] + th[2] * th[2], w, s;
            if (m2 * m2 < 5.328e-15) { w = 1.0 - 0.5 * m2; s = 1.0 - m2 / 6.0; }
            else { double mg = sqrt(m2); w = cos(mg); s = sin(mg) / mg; }
            E.rotQ[0] = w; E.rotQ[1] = th[0] * s; E.rotQ[2] = th[1] * s; E.rotQ[3] = th[2] * s;
            E.hasRotQ = 1;
        }
    }

    // importer temperature (VX_Sim.cpp:146-147 + VX_Environment.cpp:327-346)
    H.tInit = 0.0f;
    if (d->temp_enabled) {
        float cur;
        if (d->vary_temp_enabled) cur = d->temp_period == 0 ? 0.0f : (float)(d->temp_base + d->temp_amplitude * sin(2 * 3.1415926 / d->temp_period * 0.0));
        else cur = (float)(d->temp_base + d->temp_amplitude);
        H.tInit = (float)(cur - d->temp_base);
    }
    const bool vary = d->vary_temp_enabled && d->temp_period > 0;
    H.tfac0.resize(NV);
    for (int v = 0; v < NV; v++) {
        const Mat& m = mats[H.vmat[v]];
        float T = H.tInit;
        // VX3 temperature update at t = 0 (VX3_VoxelyzeKernel.cu:625-650)
        if (vary && !(m.thermalOn > 0.0) && !m.fixed) {
            double Td = d->temp_amplitude * sin((double)(2 * 3.1415926f) * (0.0 / d->temp_period + H.phase[v]));
            if (!d->enable_expansion && Td > 0) Td = 0;
            T = (float)Td;
        }
        H.tfac0[v] = 1 + T * m.cte;
    }

    // poisson, initial center of mass (VX3_VoxelyzeKernel.cu:477-493)
    H.anyPoisson = false;
    for (int i = 0; i < NM; i++) if (mats[i].nu != 0) H.anyPoisson = true;
    {
        double tot = 0, s[3] = {0, 0, 0};
        for (int v = 0; v < NV; v++) {
            const VoxMatDev& m = H.vmats[H.vmat[v]];
            if (!m.isMeasured) continue;
            double mass = m.mass;
            for (int k = 0; k < 3; k++) s[k] += (d->ijk[3 * v + k] * d->voxel_size) * mass;
            tot += mass;
        }
        for (int k = 0; k < 3; k++) H.com0[k] = tot == 0 ? 0.0 : (1.0 / tot) * s[k];
    }
    H.flags = 0;
    if (vary) H.flags |= SC_VARY;
    if (d->enable_expansion) H.flags |= SC_EXPANSION;
    if (H.anyPoisson) H.flags |= SC_POISSON;
    if (d->floor_enabled) H.flags |= SC_FLOOR;
    if (d->collisions_enabled) H.flags |= SC_COLLISIONS;
    H.tempAmp = d->temp_amplitude;
    H.tempPeriod = d->temp_period;
    H.dtFrac = d->dt_frac;
    H.dtFixed = d->dt_fixed;
    H.maxSteps = d->max_steps;
    H.collisions = d->collisions_enabled != 0;
    H.boundingRadius = d->bounding_radius;
    H.watchDistance = d->watch_distance;
    if (d->n_stop_tokens > 0) H.stopTok.assign(d->stop_tokens, d->stop_tokens + d->n_stop_tokens);
    if (d->n_fitness_tokens > 0) H.fitTok.assign(d->fitness_tokens, d->fitness_tokens + d->n_fitness_tokens);
    H.saveAll = d->save_position_of_all_voxels;
    if (d->enable_cilia) {
        H.vmCilia.resize(d->n_materials);
        H.vmCiliaOn.resize(d->n_materials);
        bool any = false;
        for (int i = 0; i < d->n_materials; i++) {
            H.vmCilia[i] = d->materials[i].cilia;
            H.vmCiliaOn[i] = d->materials[i].cilia_on_after;
            any |= d->materials[i].cilia != 0;
        }
        if (any) {
            H.cilia.assign(6 * (size_t)H.nv, 0.0);
            for (int v = 0; v < H.nv; v++)
                for (int k = 0; k < 3; k++) {
                    if (d->base_cilia_force) H.cilia[6 * (size_t)v + k] = d->base_cilia_force[3 * (size_t)v + k];
                    if (d->shift_cilia_force) H.cilia[6 * (size_t)v + 3 + k] = d->shift_cilia_force[3 * (size_t)v + k];
                }
        }
    }
    for (int k = 0; k < 3; k++) {
        const int n = d->n_force_field_tokens[k];
        if (n < 0 || n > VX3H_MAX_TOKENS || (n > 0 && !d->force_field_tokens[k])) {
            err = "bad force-field program";
            return false;
        }
        if (n > 0) H.ffTok[k].assign(d->force_field_tokens[k], d->force_field_tokens[k] + n);
    }
    if (d->enable_attach)
        for (int k = 0; k < 5; k++) {
            const int n = d->n_attach_tokens[k];
            if (n < 0 || n > VX3H_MAX_TOKENS || (n > 0 && !d->attach_tokens[k])) {
                err = "bad attach-condition program";
                return false;
            }
            if (n > 0) H.atTok[k].assign(d->attach_tokens[k], d->attach_tokens[k] + n);
        }
    return true;
}

double evalMathTree(const vx3h_token* buf, int n, double x, double y, double z, double hit, double t, double angle,
                    double closeness, int ncp, int nvox) {
    if (!buf || n <= 0) return 0.0;
    std::vector<double> q(n + 1, 0.0);
    int vc = 0, pc = 0;
    for (int i = 0; i < n; i++) {
        const int op = buf[i].op;
        const double val = buf[i].value;
        switch (op) {
        case 0: return q[pc];
        case 1: q[vc] = val; break;
        case 2: q[vc] = 2.71828182845904523536; break;
        case 3: q[vc] = 3.14159265358979323846; break;
        case 4: {
            const double vars[9] = {x, y, z, hit, t, angle, closeness, (double)ncp, (double)nvox};
            int k = val < 0.5 ? 0 : (int)(val + 0.5);
            q[vc] = (k >= 0 && k < 9) ? vars[k] : 0.0;
            break;
        }
        case 5: q[vc] = q[pc + 1] + q[pc]; pc += 2; break;
        case 6: q[vc] = q[pc + 1] - q[pc]; pc += 2; break;
        case 7: q[vc] = q[pc + 1] * q[pc]; pc += 2; break;
        case 8: q[vc] = q[pc + 1] / q[pc]; pc += 2; break;
        case 9: q[vc] = pow(q[pc + 1], q[pc]); pc += 2; break;
        case 10: q[vc] = sqrt(q[pc]); pc++; break;
        case 11: q[vc] = sin(q[pc]); pc++; break;
        case 12: q[vc] = cos(q[pc]); pc++; break;
        case 13: q[vc] = tan(q[pc]); pc++; break;
        case 14: q[vc] = atan(q[pc]); pc++; break;
        case 15: q[vc] = log(q[pc]); pc++; break;
        case 16: q[vc] = rint(q[pc]); pc++; break;
        case 17: q[vc] = fabs(q[pc]); pc++; break;
        case 18: q[vc] = !q[pc]; pc++; break;
        case 19: q[vc] = q[pc + 1] > q[pc]; pc += 2; break;
        case 20: q[vc] = q[pc + 1] < q[pc]; pc += 2; break;
        case 21: q[vc] = q[pc + 1] && q[pc]; pc += 2; break;
        case 22: q[vc] = q[pc + 1] || q[pc]; pc += 2; break;
        case 23: q[vc] = 0.5 * erfc(-q[pc] / sqrt(2.0)); pc++; break;
        default: return -1;
        }
        if (pc > vc) return -1;
        vc++;
    }
    return q[pc];
}

}  // namespace vx3
