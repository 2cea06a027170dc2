/*
 * vx_oracle.cpp -- CPU oracle of the Voxelyze3 physics path.
 *
 * TEST INFRASTRUCTURE: this file restates the reference CPU engine so that the
 * HIP engine can be checked against it.  It is never linked into the product.
 *
 * Precision and operation order follow the reference exactly (positions,
 * momenta and orientations in double, strains/stresses/material constants in
 * float), so that, compiled with g++ -O2 (no FMA contraction on x86-64), it is
 * bit-identical to the reference sources built by oracle/Makefile.
 *
 * Reference files restated (paths relative to the gpuVoxels tree):
 *   Vec3D / Quat3D       Voxelyze3/src/old/Vec3D.h:31-113, Quat3D.h:21-210
 *   materials            Voxelyze3/src/old/VX_Material.cpp:15-549,
 *                        VX_MaterialVoxel.cpp:17-86, VX_MaterialLink.cpp:15-141
 *   link beam element    Voxelyze3/src/old/VX_Link.cpp:21-259, VX_Link.h:115-120
 *   voxel integrator     Voxelyze3/src/old/VX_Voxel.cpp:23-417, VX_Voxel.h:88-130
 *   collisions           Voxelyze3/src/old/VX_Collision.cpp:15-56
 *   externals            Voxelyze3/src/old/VX_External.cpp:14-113
 *   engine               Voxelyze3/src/old/Voxelyze.cpp:251-751
 *   import               Voxelyze3/src/VXA/VX_Sim.cpp:59-153,368-483
 *   VX3 driver           Voxelyze3/src/VX3/VX3_VoxelyzeKernel.cu:184-359,477-534,625-650,
 *                        VX3_SimulationManager.cu:52-117
 *   MathTree             Voxelyze3/src/Utils/VX3_MathTree.h:50-192
 *   VX3 extras           cilia VX3_VoxelyzeKernel.cu:846-859, VX3_Voxel.cu:376; force field
 *                        VX3_Voxel.cu:178-184; detach VX3_VoxelyzeKernel.cu:946-967; voxel
 *                        removal and reset :336-350,365-397; signals VX3_Voxel.cu:271-348;
 *                        attach VX3_VoxelyzeKernel.cu:608-621,651-680,724-843, VX3_Link.cu:30-55,206-213
 *   VoxCAD stop family   vx_voxcad_stop.h (VoxCAD/src/VX1/VX_Sim.cpp:862-891,1015-1160,1833-1837)
 *
 * How each part is pinned: the CPU engine and importer bit-exactly against the
 * reference's own sources compiled in place (oracle/_ref); the VX3 driver and the
 * VX3 extras, which exist only as GPU code in the reference, by closed forms and
 * brute-force restatements in tests/test_oracle_kat.py (phase temperature, dt, CoM,
 * fitness, stop counts; cilia momentum, rotation and surface rule; force fields;
 * detach; removal and reset; signal hops, pacemaker period and decay, cilia
 * coupling; attach: which pairs link, worked out by hand from the five-link rule,
 * the sticky-material and condition gates, no force in the step a link is made and
 * exactly 1/100 of it at the next).  Attach takes pairs one after another in the
 * reference's pair order where its threads race (VX3_VoxelyzeKernel.cu:833-843): that
 * order is this oracle's definition, as for signals.  Signals follow a deterministic two-phase order where the reference's
 * threads race (VX3_Voxel.cu:323-348): that order is this oracle's definition, so
 * signal parity with the reference is unpinned beyond the single-hop behaviour.
 */
#include "vx_oracle.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
/* <math.h> (C++ wrapper) puts the float overloads of sqrt/pow/... in the global
 * namespace, as in the reference translation units (Vec3D.h includes it), so a
 * float argument selects single precision exactly as there. */
#include <math.h>
#include <cstdint>
#include <cstring>
#include <functional>
#include <list>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "vx_mathtree.h"
#include "vx_voxcad_stop.h"

namespace vxo {

/* ------------------------------------------------------------------ math */
#define O_PI 3.14159265358979
#define O_DBL_EPSILONx24 5.328e-15
#define O_DISCARD_ANGLE_RAD 1e-7
#define O_SMALL_ANGLE_RAD 1.732e-2
#define O_SLTHRESH_ACOS2SQRT 2.4e-3

static const float HYSTERESIS_FACTOR = 1.2f;   // old/types.h:57
static const float SA_BOND_BEND_RAD = 0.05f;   // old/types.h:58
static const float SA_BOND_EXT_PERC = 0.50f;   // old/types.h:59

template <typename T> struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    template <typename U> V3(const V3<U>& s) : x((T)s.x), y((T)s.y), z((T)s.z) {}
    V3 operator+(const V3& v) const { return V3(x + v.x, y + v.y, z + v.z); }
    V3 operator-(const V3& v) const { return V3(x - v.x, y - v.y, z - v.z); }
    template <typename U> V3 operator-(const V3<U>& v) const { return V3(x - v.x, y - v.y, z - v.z); }
    V3 operator-() const { return V3(-x, -y, -z); }
    template <typename U> V3 operator*(const U& f) const { return V3(f * x, f * y, f * z); }
    V3 operator/(const T& f) const { T inv = (T)1.0 / f; return V3(inv * x, inv * y, inv * z); }
    template <typename U> V3& operator+=(const V3<U>& s) { x += s.x; y += s.y; z += s.z; return *this; }
    template <typename U> V3& operator-=(const V3<U>& s) { x -= s.x; y -= s.y; z -= s.z; return *this; }
    bool operator==(const V3& v) const { return x == v.x && y == v.y && z == v.z; }
    T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T Length() const { return sqrt(x * x + y * y + z * z); }
    T Length2() const { return x * x + y * y + z * z; }
    T Dot(const V3& v) const { return x * v.x + y * v.y + z * v.z; }
    T Dist(const V3& v) const { return sqrt(Dist2(v)); }
    T Dist2(const V3& v) const { return (v.x - x) * (v.x - x) + (v.y - y) * (v.y - y) + (v.z - z) * (v.z - z); }
    void NormalizeFast() { T l = sqrt(x * x + y * y + z * z); if (l > 0) { T li = 1.0 / l; x *= li; y *= li; z *= li; } }
    V3 Normalized() const { T l = sqrt(x * x + y * y + z * z); return (l > 0 ? (*this) / l : (*this)); }
    V3 Scale(const V3& v) const { return V3(x * v.x, y * v.y, z * v.z); }
};
template <typename U, typename T> V3<T> operator*(const U f, const V3<T>& v) { return v * f; }
typedef V3<double> Vd;
typedef V3<float> Vf;

template <typename T> struct Q4 {
    T w, x, y, z;
    Q4() : w(1), x(0), y(0), z(0) {}
    Q4(T a, T b, T c, T d) : w(a), x(b), y(c), z(d) {}
    explicit Q4(const V3<T>& rv) { FromRotationVector(rv); }
    Q4 operator*(const Q4& f) const {
        return Q4(w * f.w - x * f.x - y * f.y - z * f.z, w * f.x + x * f.w + y * f.z - z * f.y,
                  w * f.y - x * f.z + y * f.w + z * f.x, w * f.z + x * f.y - y * f.x + z * f.w);
    }
    Q4 Conjugate() const { return Q4(w, -x, -y, -z); }
    V3<T> ToRotationVector() const {
        if (w >= 1.0 || w <= -1.0) return V3<T>(0, 0, 0);
        T squareLength = 1.0 - w * w;
        if (squareLength < O_SLTHRESH_ACOS2SQRT) return V3<T>(x, y, z) * 2.0 * sqrt((2 - 2 * w) / squareLength);
        else return V3<T>(x, y, z) * 2.0 * acos(w) / sqrt(squareLength);
    }
    void FromRotationVector(const V3<T>& VecIn) {
        V3<T> theta = VecIn / 2;
        T s, thetaMag2 = theta.Length2();
        if (thetaMag2 * thetaMag2 < O_DBL_EPSILONx24) {
            w = 1.0 - 0.5 * thetaMag2;
            s = 1.0 - thetaMag2 / 6.0;
        } else {
            T thetaMag = sqrt(thetaMag2);
            w = cos(thetaMag);
            s = sin(thetaMag) / thetaMag;
        }
        x = theta.x * s;
        y = theta.y * s;
        z = theta.z * s;
    }
    void FromAngleToPosX(const V3<T>& RotateFrom) {
        if (V3<T>(0, 0, 0) == RotateFrom) return;
        T YoverX = RotateFrom.y / RotateFrom.x;
        T ZoverX = RotateFrom.z / RotateFrom.x;
        if (YoverX < O_SMALL_ANGLE_RAD && YoverX > -O_SMALL_ANGLE_RAD && ZoverX < O_SMALL_ANGLE_RAD && ZoverX > -O_SMALL_ANGLE_RAD) {
            x = 0; y = 0.5 * ZoverX; z = -0.5 * YoverX;
            w = 1 + 0.5 * (-y * y - z * z);
            return;
        }
        V3<double> RotFromNorm = RotateFrom;
        RotFromNorm.NormalizeFast();
        T theta = acos(RotFromNorm.x);
        if (theta > O_PI - O_DISCARD_ANGLE_RAD) { w = 0; x = 0; y = 1; z = 0; return; }
        const T AxisMagInv = 1.0 / sqrt(RotFromNorm.z * RotFromNorm.z + RotFromNorm.y * RotFromNorm.y);
        const T a = 0.5 * theta;
        const T s = sin(a);
        w = cos(a); x = 0; y = RotFromNorm.z * AxisMagInv * s; z = -RotFromNorm.y * AxisMagInv * s;
    }
    V3<T> RotateVec3D(const V3<T>& f) const {
        T fx = f.x, fy = f.y, fz = f.z;
        T tw = fx * x + fy * y + fz * z;
        T tx = fx * w - fy * z + fz * y;
        T ty = fx * z + fy * w - fz * x;
        T tz = -fx * y + fy * x + fz * w;
        return V3<T>(w * tx + x * tw + y * tz - z * ty, w * ty - x * tz + y * tw + z * tx, w * tz + x * ty - y * tx + z * tw);
    }
    V3<float> RotateVec3Df(const V3<float>& f) const {   // RotateVec3D<U=float> (Quat3D.h:183-190)
        float fx = f.x, fy = f.y, fz = f.z;
        float tw = (float)(fx * x + fy * y + fz * z);
        float tx = (float)(fx * w - fy * z + fz * y);
        float ty = (float)(fx * z + fy * w - fz * x);
        float tz = (float)(-fx * y + fy * x + fz * w);
        return V3<float>((float)(w * tx + x * tw + y * tz - z * ty), (float)(w * ty - x * tz + y * tw + z * tx),
                         (float)(w * tz + x * ty - y * tx + z * tw));
    }
    V3<T> RotateVec3DInv(const V3<T>& f) const {
        T fx = f.x, fy = f.y, fz = f.z;
        T tw = x * fx + y * fy + z * fz;
        T tx = w * fx - y * fz + z * fy;
        T ty = w * fy + x * fz - z * fx;
        T tz = w * fz - x * fy + y * fx;
        return V3<T>(tw * x + tx * w + ty * z - tz * y, tw * y - tx * z + ty * w + tz * x, tw * z + tx * y - ty * x + tz * w);
    }
};
typedef Q4<double> Qd;

/* ------------------------------------------------------------------ materials */
struct Material {
    int matid = 0;
    bool fixed = false;
    int isMeasured = 1;
    double thermalOnAfter = 0.0;
    bool linear = true;
    float E = 0, sigmaYield = 0, sigmaFail = 0, epsilonYield = 0, epsilonFail = 0;
    std::vector<float> strainData, stressData;
    float nu = 0, rho = 0, alphaCTE = 0, muStatic = 0, muKinetic = 0;
    float zetaInternal = 0, zetaGlobal = 0, zetaCollision = 0;
    Vd extScale;
    float _eHat = 0;
    std::vector<Material*> dependentMaterials;
    virtual ~Material() {}
    virtual bool updateAll() { return false; }
    // VX_Material.cpp:542-549
    virtual bool updateDerived() {
        _eHat = E / ((1 - 2 * nu) * (1 + nu));
        for (Material* m : dependentMaterials) m->updateAll();
        return true;
    }
    // VX_Material.cpp:71-90
    void clear() {
        nu = 0.0f; rho = 1.0f; alphaCTE = 0.0f; muStatic = 0.0f; muKinetic = 0.0f;
        zetaInternal = 1.0f; zetaGlobal = 0.0f; zetaCollision = 0.0f;
        extScale = Vd(1.0, 1.0, 1.0);
        setModelLinear(1.0);
        updateDerived();
    }
    bool isFailed(float strain) const { return epsilonFail != -1.0f && strain > epsilonFail; }
    bool isYielded(float strain) const { return epsilonYield != -1.0f && strain > epsilonYield; }
    bool isXyzIndependent() const { return nu == 0.0f; }
    // VX_Material.cpp:182-212
    float stress(float strain, float transverseStrainSum = 0.0f, bool forceLinear = false) {
        if (isFailed(strain)) return 0.0f;
        if (strain <= strainData[1] || linear || forceLinear) {
            if (nu == 0.0f) return E * strain;
            else return _eHat * ((1 - nu) * strain + nu * transverseStrainSum);
        }
        int DataCount = (int)strainData.size();
        for (int i = 2; i < DataCount; i++) {
            if (strain <= strainData[i] || i == DataCount - 1) {
                float Perc = (strain - strainData[i - 1]) / (strainData[i] - strainData[i - 1]);
                float basicStress = stressData[i - 1] + Perc * (stressData[i] - stressData[i - 1]);
                if (nu == 0.0f) return basicStress;
                else {
                    float modulus = (stressData[i] - stressData[i - 1]) / (strainData[i] - strainData[i - 1]);
                    float modulusHat = modulus / ((1 - 2 * nu) * (1 + nu));
                    float effectiveStrain = basicStress / modulus;
                    float effectiveTransverseStrainSum = transverseStrainSum * (effectiveStrain / strain);
                    return modulusHat * ((1 - nu) * effectiveStrain + nu * effectiveTransverseStrainSum);
                }
            }
        }
        return 0.0f;
    }
    // VX_Material.cpp:214-228
    float strainOf(float stress) {
        if (stress <= stressData[1] || linear) return stress / E;
        int DataCount = (int)strainData.size();
        for (int i = 2; i < DataCount; i++) {
            if (stress <= stressData[i] || i == DataCount - 1) {
                float Perc = (stress - stressData[i - 1]) / (stressData[i] - stressData[i - 1]);
                return strainData[i - 1] + Perc * (strainData[i] - strainData[i - 1]);
            }
        }
        return 0.0f;
    }
    // VX_Material.cpp:231-243
    float modulus(float strain) {
        if (isFailed(strain)) return 0.0f;
        if (strain <= strainData[1] || linear) return E;
        int DataCount = (int)strainData.size();
        for (int i = 2; i < DataCount; i++) {
            if (strain <= strainData[i] || i == DataCount - 1) return (stressData[i] - stressData[i - 1]) / (strainData[i] - strainData[i - 1]);
        }
        return 0.0f;
    }
    // VX_Material.cpp:300-367
    bool setModel(int dataPointCount, const float* pStrainValues, const float* pStressValues) {
        if (*pStrainValues == 0 && *pStressValues == 0) { pStrainValues++; pStressValues++; dataPointCount--; }
        if (dataPointCount <= 0) return false;
        if (*pStrainValues <= 0 || *pStressValues <= 0) return false;
        std::vector<float> tmpStrainData, tmpStressData;
        tmpStrainData.push_back(0); tmpStressData.push_back(0);
        float sweepStrain = 0.0f, sweepStress = 0.0f;
        for (int i = 0; i < dataPointCount; i++) {
            float thisStrain = *(pStrainValues + i);
            float thisStress = *(pStressValues + i);
            if (thisStrain <= sweepStrain) return false;
            if (i > 0 && (thisStress - sweepStress) / (thisStrain - sweepStrain) > tmpStressData[0] / tmpStrainData[0]) return false;
            sweepStrain = thisStrain; sweepStress = thisStress;
            tmpStrainData.push_back(thisStrain); tmpStressData.push_back(thisStress);
        }
        strainData = tmpStrainData; stressData = tmpStressData;
        E = stressData[1] / strainData[1];
        sigmaFail = stressData[stressData.size() - 1];
        epsilonFail = strainData[strainData.size() - 1];
        linear = (dataPointCount == 1);
        if (dataPointCount == 1 || dataPointCount == 2) { sigmaYield = stressData[1]; epsilonYield = strainData[1]; }
        else setYieldFromData();
        return updateDerived();
    }
    // VX_Material.cpp:372-401
    bool setModelLinear(float youngsModulus, float failureStress = -1) {
        if (youngsModulus <= 0) return false;
        if (failureStress != -1.0f && failureStress <= 0) return false;
        float tmpfailureStress = failureStress;
        if (tmpfailureStress == -1) tmpfailureStress = 1000000;
        float tmpfailStrain = tmpfailureStress / youngsModulus;
        strainData.clear(); stressData.clear();
        strainData.push_back(0); stressData.push_back(0);
        strainData.push_back(tmpfailStrain); stressData.push_back(tmpfailureStress);
        linear = true; E = youngsModulus;
        sigmaYield = failureStress; sigmaFail = failureStress;
        epsilonYield = (failureStress == -1) ? -1 : tmpfailStrain;
        epsilonFail = (failureStress == -1) ? -1 : tmpfailStrain;
        return updateDerived();
    }
    // VX_Material.cpp:406-451
    bool setModelBilinear(float youngsModulus, float plasticModulus, float yieldStress, float failureStress = -1) {
        if (youngsModulus <= 0) return false;
        if (plasticModulus <= 0 || plasticModulus >= youngsModulus) return false;
        if (yieldStress <= 0) return false;
        if (failureStress != -1.0f && failureStress <= yieldStress) return false;
        float yieldStrain = yieldStress / youngsModulus;
        float tmpfailureStress = failureStress;
        if (tmpfailureStress == -1) tmpfailureStress = 3 * yieldStress;
        float tM = plasticModulus;
        float tB = yieldStress - tM * yieldStrain;
        float tmpfailStrain = (tmpfailureStress - tB) / tM;
        strainData.clear(); strainData.push_back(0); strainData.push_back(yieldStrain); strainData.push_back(tmpfailStrain);
        stressData.clear(); stressData.push_back(0); stressData.push_back(yieldStress); stressData.push_back(tmpfailureStress);
        linear = false; E = youngsModulus;
        sigmaYield = yieldStress; sigmaFail = failureStress;
        epsilonYield = yieldStrain;
        epsilonFail = failureStress == -1.0f ? -1.0f : tmpfailStrain;
        return updateDerived();
    }
    // VX_Material.cpp:454-487
    bool setYieldFromData(float percentStrainOffset = 0.2) {
        sigmaYield = -1.0f; epsilonYield = -1.0f;
        float oM = E;
        float oB = (-percentStrainOffset / 100 * oM);
        int dataPoints = (int)strainData.size() - 1;
        for (int i = 1; i < dataPoints - 1; i++) {
            float x1 = strainData[i], x2 = strainData[i + 1], y1 = stressData[i], y2 = stressData[i + 1];
            float tM = (y2 - y1) / (x2 - x1);
            float tB = y1 - tM * x1;
            if (oM != tM) {
                float xIntersect = (tB - oB) / (oM - tM);
                if (xIntersect > x1 && xIntersect < x2) {
                    float percentBetweenPoints = (xIntersect - x1) / (x2 - x1);
                    sigmaYield = y1 + percentBetweenPoints * (y2 - y1);
                    epsilonYield = xIntersect;
                    return true;
                }
            }
        }
        sigmaYield = sigmaFail; epsilonYield = epsilonFail;
        return false;
    }
    // VX_Material.cpp:489-540
    void setPoissonsRatio(float p) { if (p < 0) p = 0; if (p >= 0.5) p = 0.5 - FLT_EPSILON * 2; nu = p; updateDerived(); }
    void setDensity(float d) { if (d <= 0) d = FLT_MIN; rho = d; updateDerived(); }
    void setStaticFriction(float v) { if (v <= 0) v = 0; muStatic = v; }
    void setKineticFriction(float v) { if (v <= 0) v = 0; muKinetic = v; }
    void setInternalDamping(float z) { if (z <= 0) z = 0; zetaInternal = z; }
    void setGlobalDamping(float z) { if (z <= 0) z = 0; zetaGlobal = z; }
    void setCollisionDamping(float z) { if (z <= 0) z = 0; zetaCollision = z; }
};

struct MaterialVoxel : Material {
    bool sticky = false;   // VX3 attach
    double nomSize = 0.001;
    float gravMult = 0;
    float _mass = 0, _massInverse = 0, _sqrtMass = 0, _firstMoment = 0, _momentInertia = 0, _momentInertiaInverse = 0;
    float _2xSqMxExS = 0, _2xSqIxExSxSxS = 0;
    // CVX_MaterialVoxel(E, rho, size) -> CVX_Material(E, rho) (VX_Material.cpp:15-22), initialize
    MaterialVoxel(float E_ = 1e6f, float density = 1e3f, double nominalSize = 0.001) {
        clear();
        rho = density;
        setModelLinear(E_);
        Material::updateDerived();
        nomSize = nominalSize; gravMult = 0.0f; updateDerived();
    }
    // VX_MaterialVoxel.cpp:57-79
    bool updateDerived() override {
        Material::updateDerived();
        double volume = nomSize * nomSize * nomSize;
        _mass = (float)(volume * rho);
        _momentInertia = (float)(_mass * nomSize * nomSize / 6.0f);
        _firstMoment = (float)(_mass * nomSize / 2.0f);
        if (volume == 0 || _mass == 0 || _momentInertia == 0) {
            _massInverse = _sqrtMass = _momentInertiaInverse = _2xSqMxExS = _2xSqIxExSxSxS = 0.0f;
            return false;
        }
        _massInverse = 1.0f / _mass;
        _sqrtMass = sqrt(_mass);
        _momentInertiaInverse = 1.0f / _momentInertia;
        _2xSqMxExS = (float)(2.0f * sqrt(_mass * E * nomSize));
        _2xSqIxExSxSxS = (float)(2.0f * sqrt(_momentInertia * E * nomSize * nomSize * nomSize));
        return true;
    }
    bool setNominalSize(double size) { if (size <= 0) size = FLT_MIN; nomSize = size; return updateDerived(); }
    Vd size() const { return extScale * nomSize; }   // nomSize*extScale  (VX_MaterialVoxel.h:34)
    float globalDampingTranslateC() const { return zetaGlobal * _2xSqMxExS; }
    float globalDampingRotateC() const { return zetaGlobal * _2xSqIxExSxSxS; }
    float collisionDampingTranslateC() const { return zetaCollision * _2xSqMxExS; }
    float penetrationStiffness() const { return (float)(2 * E * nomSize); }
    float gravityForce() const { return -_mass * 9.80665f * gravMult; }
};

struct MaterialLink : MaterialVoxel {
    MaterialVoxel *vox1Mat, *vox2Mat;
    float _a1 = 0, _a2 = 0, _b1 = 0, _b2 = 0, _b3 = 0, _sqA1 = 0, _sqA2xIp = 0, _sqB1 = 0, _sqB2xFMp = 0, _sqB3xIp = 0;
    MaterialLink(MaterialVoxel* m1, MaterialVoxel* m2) : MaterialVoxel() { vox1Mat = m1; vox2Mat = m2; updateAll(); }
    // VX_MaterialLink.cpp:45-119
    bool updateAll() override {
        nomSize = 0.5 * (vox1Mat->nomSize + vox2Mat->nomSize);
        rho = 0.5f * (vox1Mat->rho + vox2Mat->rho);
        alphaCTE = 0.5f * (vox1Mat->alphaCTE + vox2Mat->alphaCTE);
        muStatic = 0.5f * (vox1Mat->muStatic + vox2Mat->muStatic);
        muKinetic = 0.5f * (vox1Mat->muKinetic + vox2Mat->muKinetic);
        zetaInternal = 0.5f * (vox1Mat->zetaInternal + vox2Mat->zetaInternal);
        zetaGlobal = 0.5f * (vox1Mat->zetaGlobal + vox2Mat->zetaGlobal);
        zetaCollision = 0.5f * (vox1Mat->zetaCollision + vox2Mat->zetaCollision);
        extScale = Vd(1.0, 1.0, 1.0);
        float stressFail = -1.0f, f1 = vox1Mat->sigmaFail, f2 = vox2Mat->sigmaFail;
        if (f1 == -1.0f) stressFail = f2;
        else if (f2 == -1.0f) stressFail = f1;
        else stressFail = f1 < f2 ? f1 : f2;
        if (vox1Mat->linear && vox2Mat->linear) setModelLinear(2.0f * vox1Mat->E * vox2Mat->E / (vox1Mat->E + vox2Mat->E), stressFail);
        else {
            std::vector<float> newStressValues, newStrainValues;
            newStressValues.push_back(0.0f); newStrainValues.push_back(0.0f);
            int dataIt1 = 1, dataIt2 = 1;
            while (dataIt1 < (int)vox1Mat->strainData.size() && dataIt2 < (int)vox2Mat->strainData.size()) {
                float strain = FLT_MAX;
                if (dataIt1 < (int)vox1Mat->strainData.size()) strain = vox1Mat->strainData[dataIt1];
                if (dataIt2 < (int)vox2Mat->strainData.size() && vox2Mat->strainData[dataIt2] < strain) strain = vox2Mat->strainData[dataIt2];
                if (strain == vox1Mat->strainData[dataIt1]) dataIt1++;
                if (strain == vox2Mat->strainData[dataIt2]) dataIt2++;
                float modulus1 = vox1Mat->modulus(strain - FLT_EPSILON);
                float modulus2 = vox2Mat->modulus(strain - FLT_EPSILON);
                float thisModulus = 2.0f * modulus1 * modulus2 / (modulus1 + modulus2);
                int lastDataIndex = (int)newStrainValues.size() - 1;
                newStrainValues.push_back(strain);
                newStressValues.push_back(newStressValues[lastDataIndex] + thisModulus * (strain - newStrainValues[lastDataIndex]));
            }
            setModel((int)newStrainValues.size(), &newStrainValues[0], &newStressValues[0]);
            sigmaFail = stressFail;
            epsilonFail = stressFail == -1.0f ? -1.0f : strainOf(stressFail);
        }
        if (vox1Mat->nu == 0 && vox2Mat->nu == 0) nu = 0;
        else {
            float tmpEHat = 2 * vox1Mat->_eHat * vox2Mat->_eHat / (vox1Mat->_eHat + vox2Mat->_eHat);
            float tmpE = E;
            float c2 = (tmpEHat - tmpE) / (2 * tmpEHat) + 0.0625;
            nu = sqrt(c2) - 0.25;
        }
        return updateDerived();
    }
    // VX_MaterialLink.cpp:121-141
    bool updateDerived() override {
        MaterialVoxel::updateDerived();
        float L = (float)nomSize;
        _a1 = E * L;
        _a2 = E * L * L * L / (12.0f * (1 + nu));
        _b1 = E * L;
        _b2 = E * L * L / 2.0f;
        _b3 = E * L * L * L / 6.0f;
        _sqA1 = sqrt(_a1);
        _sqA2xIp = sqrt(_a2 * L * L / 6.0f);
        _sqB1 = sqrt(_b1);
        _sqB2xFMp = sqrt(_b2 * L / 2.0f);
        _sqB3xIp = sqrt(_b3 * L * L / 6.0f);
        return true;
    }
};

/* ------------------------------------------------------------------ externals */
enum { X_TRANSLATE = 1, Y_TRANSLATE = 2, Z_TRANSLATE = 4, X_ROTATE = 8, Y_ROTATE = 16, Z_ROTATE = 32 };
struct External {
    unsigned char dofFixed = 0;
    Vf extForce, extMoment;
    Vd extTranslation, extRotation;
    bool hasRotQ = false;
    Qd rotQ;
    bool isFixed(int dof) const { return (dof & dofFixed) ? true : false; }
    bool isFixedAll() const { return (dofFixed & 0x3F) == 0x3F; }
    bool isFixedAllRotation() const { return isFixed(X_ROTATE) && isFixed(Y_ROTATE) && isFixed(Z_ROTATE); }
    bool isFixedAnyRotation() const { return isFixed(X_ROTATE) || isFixed(Y_ROTATE) || isFixed(Z_ROTATE); }
    Qd rotationQuat() const { return hasRotQ ? rotQ : Qd(); }
    // VX_External.cpp:102-112
    void rotationChanged() {
        if (!(extRotation == Vd())) { hasRotQ = true; rotQ = Qd(extRotation); }
        else if (hasRotQ) rotQ = Qd();
    }
    void setFixed(bool tx, bool ty, bool tz, bool rx, bool ry, bool rz) {   // VX_External.cpp:39-43
        dofFixed = (tx ? 1 : 0) | (ty ? 2 : 0) | (tz ? 4 : 0) | (rx ? 8 : 0) | (ry ? 16 : 0) | (rz ? 32 : 0);
        extTranslation = extRotation = Vd();
    }
    void setDisplacement(int dof, double d = 0.0) {                          // VX_External.cpp:45-58
        dofFixed |= dof;
        if (d != 0.0f) {
            if (dof & X_TRANSLATE) extTranslation.x = d;
            if (dof & Y_TRANSLATE) extTranslation.y = d;
            if (dof & Z_TRANSLATE) extTranslation.z = d;
            if (dof & X_ROTATE) extRotation.x = d;
            if (dof & Y_ROTATE) extRotation.y = d;
            if (dof & Z_ROTATE) extRotation.z = d;
        }
        rotationChanged();
    }
    void setDisplacementAll(const Vd& t, const Vd& r) { dofFixed |= 0x3F; extTranslation = t; extRotation = r; rotationChanged(); }
    void clearDisplacementAll() { dofFixed &= ~0x3F; extTranslation = Vd(); extRotation = Vd(); rotationChanged(); }
};

/* ------------------------------------------------------------------ voxel / link */
enum { SURFACE = 1 << 1, FLOOR_ENABLED = 1 << 2, FLOOR_STATIC_FRICTION = 1 << 3, COLLISIONS_ENABLED = 1 << 5 };
enum { X_POS = 0, X_NEG = 1, Y_POS = 2, Y_NEG = 3, Z_POS = 4, Z_NEG = 5 };
enum { X_AXIS = 0, Y_AXIS = 1, Z_AXIS = 2 };
enum { LOCAL_VELOCITY_VALID = 1 };

struct Voxel;
struct Collision;

struct Link {
    Voxel *pVNeg = nullptr, *pVPos = nullptr;
    bool gone = false;   // VX3 detach: no longer updated, no longer in its voxels' links
    bool fresh = false;  // VX3 attach: made this step, first updated at the next step's link pass
    int isNewLink = 0;   // VX3 attach: steps left carrying 1/100 of its force (VX3_Link.cu:206-213)
    Vd forceNeg, forcePos, momentNeg, momentPos;
    float strain = 0, maxStrain = 0, strainOffset = 0;
    int boolStates = 0;
    int axis = 0;
    MaterialLink* mat = nullptr;
    float strainRatio = 0;
    Vd pos2, angle1v, angle2v;
    Qd angle1, angle2;
    bool smallAngle = true;
    double currentRestLength = 0;
    float currentTransverseArea = 0, currentTransverseStrainSum = 0;
    float _stress = 0;

    bool valid = false;
    void init(Voxel* v1, Voxel* v2, MaterialLink* m);
    void reset();
    template <typename T> V3<T> toAxisX(const V3<T>& v) const {
        switch (axis) { case Y_AXIS: return V3<T>(v.y, -v.x, v.z); case Z_AXIS: return V3<T>(v.z, v.y, -v.x); default: return v; }
    }
    template <typename T> Q4<T> toAxisX(const Q4<T>& q) const {
        switch (axis) { case Y_AXIS: return Q4<T>(q.w, q.y, -q.x, q.z); case Z_AXIS: return Q4<T>(q.w, q.z, q.y, -q.x); default: return q; }
    }
    template <typename T> void toAxisOriginal(V3<T>* pV) const {
        switch (axis) {
        case Y_AXIS: { T tmp = pV->y; pV->y = pV->x; pV->x = -tmp; break; }
        case Z_AXIS: { T tmp = pV->z; pV->z = pV->x; pV->x = -tmp; break; }
        default: break;
        }
    }
    Vd force(bool positiveEnd) const { return positiveEnd ? forcePos : forceNeg; }
    Vd moment(bool positiveEnd) const { return positiveEnd ? momentPos : momentNeg; }
    float axialStrain() const { return strain; }
    float axialStrain(bool positiveEnd) const {
        return positiveEnd ? 2.0f * strain * strainRatio / (1.0f + strainRatio) : 2.0f * strain / (1.0f + strainRatio);
    }
    bool isFailed() const { return mat->isFailed(maxStrain); }
    // VX_Link.cpp:243-249 (tensile, torsion, bending about z and about y)
    float strainEnergy() const {
        return forceNeg.x * forceNeg.x / (2.0f * mat->_a1) + momentNeg.x * momentNeg.x / (2.0 * mat->_a2) +
               (momentNeg.z * momentNeg.z - momentNeg.z * momentPos.z + momentPos.z * momentPos.z) / (3.0 * mat->_b3) +
               (momentNeg.y * momentNeg.y - momentNeg.y * momentPos.y + momentPos.y * momentPos.y) / (3.0 * mat->_b3);
    }
    bool isLocalVelocityValid() const { return boolStates & LOCAL_VELOCITY_VALID ? true : false; }
    void setBoolState(int flag, bool set = true) { set ? boolStates |= flag : boolStates &= ~flag; }
    void updateRestLength();
    void updateTransverseInfo();
    Qd orientLink();
    void updateForces();
    float updateStrain(float axialStrain);
    float axialStiffness();
};

struct Voxel {
    MaterialVoxel* mat = nullptr;
    int ix = 0, iy = 0, iz = 0;
    External* ext = nullptr;
    Link* links[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    Vd pos, linMom, angMom;
    Qd orient;
    // VX3 extras (GPU-only in the reference): this step's cilia force, added at the end
    // of force() (VX3_Voxel.cu:376), and force-field force, added after force()
    // (VX3_Voxel.cu:178-184); used when the simulation has either
    bool extras = false;
    bool removed = false;   // VX3 secondary experiment: frozen, without links
    bool enableAttach = true;   // VX3 attach conditions as of the last update (VX3_Voxel.h:302)
    bool target = false;    // material isTarget
    // VX3 signals (VX3_Voxel.h:305-309): local signal and its decay time, next pacemaker
    // pulse, deaf-until time, pending signal, and the signal posted this step
    double localSignal = 0, localSignaldt = 0, nextPulse = 0, inactiveUntil = 0, pendingValue = 0, pendingTime = 0;
    double postedValue = 0, postedTime = 0;
    int isPaceMaker = 0;
    double paceMakerPeriod = 0, signalValueDecay = 0.9, signalTimeDelay = 0.03, inactivePeriod = 0.03;
    // VX3_Voxel::receiveSignal (VX3_Voxel.cu:301-322)
    void receiveSignal(double value, double activeTime, bool force) {
        if (!force && inactiveUntil > activeTime) return;
        if (value < 0.1) return;
        inactiveUntil = activeTime + inactivePeriod;
        localSignal = value;
        pendingValue = value * signalValueDecay;
        if (pendingValue < 0.1) pendingValue = 0;
        pendingTime = activeTime;
    }
    // the end of VX3_Voxel::timeStep (VX3_Voxel.cu:271-276) on this voxel's own state:
    // propagateSignal posts the pending signal, packMaker, localSignalDecay
    void signalsOwn(double t) {
        postedValue = postedTime = 0;
        if (!(pendingTime > t) && !(pendingValue < 0.1)) {
            postedValue = pendingValue;
            postedTime = t + signalTimeDelay;
            pendingValue = pendingTime = 0;
            inactiveUntil = t + 2 * signalTimeDelay + inactivePeriod;
        }
        if (isPaceMaker && !(nextPulse > t)) {
            receiveSignal(100, t, true);
            nextPulse = t + paceMakerPeriod;
        }
        if (!(localSignaldt > t)) {
            if (localSignal < 0.1) localSignal = 0;
            else {
                localSignal = localSignal * 0.9;
                localSignaldt = t + 0.01;
            }
        }
    }
    Vd ciliaForce, fieldForce;
    Vd baseCilia, shiftCilia;
    int boolStates = 0;
    float temp = 0;
    Vf pStrain;
    bool poissonsStrainInvalid = true;
    float previousDt = 0;
    Vf* lastColWatchPosition = nullptr;
    std::vector<Collision*>* colWatch = nullptr;
    std::vector<Voxel*>* nearby = nullptr;
    double phaseOffset = 0;
    int index = 0;   // position in the voxel list

    ~Voxel() { delete lastColWatchPosition; delete colWatch; delete nearby; delete ext; }
    External* external() { if (!ext) ext = new External(); return ext; }
    Vd originalPosition() const { double s = mat->nomSize; return Vd(ix * s, iy * s, iz * s); }
    void haltMotion() { linMom = angMom = Vd(0, 0, 0); }
    void setFloorStaticFriction(bool a) { a ? boolStates |= FLOOR_STATIC_FRICTION : boolStates &= ~FLOOR_STATIC_FRICTION; }
    bool isFloorStaticFriction() const { return boolStates & FLOOR_STATIC_FRICTION ? true : false; }
    bool isFloorEnabled() const { return boolStates & FLOOR_ENABLED ? true : false; }
    void enableFloor(bool e) { e ? boolStates |= FLOOR_ENABLED : boolStates &= ~FLOOR_ENABLED; }
    bool isInterior() const { return (boolStates & SURFACE) ? true : false; }
    bool isSurface() const { return !isInterior(); }
    bool isCollisionsEnabled() const { return boolStates & COLLISIONS_ENABLED ? true : false; }
    // VX_Voxel.h:88-92
    Vd baseSize() const { return mat->size() * (1 + temp * mat->alphaCTE); }
    double baseSize(int ax) const { return mat->size()[ax] * (1 + temp * mat->alphaCTE); }
    double baseSizeAverage() const { Vd b = baseSize(); return (b.x + b.y + b.z) / 3.0f; }
    Vd velocity() const { return linMom * mat->_massInverse; }
    Vd angularVelocity() const { return angMom * mat->_momentInertiaInverse; }
    // VX_Voxel.h:104
    float kineticEnergy() const {
        return (float)(0.5 * (mat->_massInverse * linMom.Length2() + mat->_momentInertiaInverse * angMom.Length2()));
    }
    float floorPenetration() const { return (float)(baseSizeAverage() / 2 - mat->nomSize / 2 - pos.z); }
    float dampingMultiplier() const { return 2 * mat->_sqrtMass * mat->zetaInternal / previousDt; }
    static bool isNegative(int d) { return d % 2 == 1; }
    static int toAxis(int d) { return d / 2; }
    static int toOpposite(int d) { return d - d % 2 + (d + 1) % 2; }
    Voxel* adjacentVoxel(int d) const {
        Link* pL = links[d];
        if (pL) return pL->pVPos == this ? pL->pVNeg : pL->pVPos;
        return nullptr;
    }
    // VX_Voxel.cpp:48-57
    void reset() {
        pos = originalPosition();
        orient = Qd();
        haltMotion();
        setFloorStaticFriction(true);
        temp = 0.0f;
        previousDt = 0.0f;
        poissonsStrainInvalid = true;
    }
    void updateSurface() {
        bool interior = true;
        for (int i = 0; i < 6; i++) if (!links[i]) interior = false;
        interior ? boolStates |= SURFACE : boolStates &= ~SURFACE;
    }
    void setTemperature(float t) {
        temp = t;
        for (int i = 0; i < 6; i++) if (links[i] != nullptr) links[i]->updateRestLength();
    }
    void enableCollisions(bool en) {
        if (en) {
            if (!lastColWatchPosition) lastColWatchPosition = new Vf;
            if (!colWatch) colWatch = new std::vector<Collision*>;
            if (!nearby) nearby = new std::vector<Voxel*>;
        }
        en ? boolStates |= COLLISIONS_ENABLED : boolStates &= ~COLLISIONS_ENABLED;
    }
    // VX_Voxel.cpp:299-333
    Vf strainV(bool poissonsStrain) const {
        Vf intStrRet(0, 0, 0);
        int numBondAxis[3] = {0};
        bool tension[3] = {false};
        for (int i = 0; i < 6; i++) {
            if (links[i]) {
                int ax = toAxis(i);
                intStrRet[ax] += links[i]->axialStrain(isNegative(i));
                numBondAxis[ax]++;
            }
        }
        for (int i = 0; i < 3; i++) {
            if (numBondAxis[i] == 2) intStrRet[i] *= 0.5f;
            if (poissonsStrain) {
                tension[i] = ((numBondAxis[i] == 2) ||
                              (ext && (numBondAxis[i] == 1 && (ext->isFixed(1 << i) || ext->extForce[i] != 0))));
            }
        }
        if (poissonsStrain) {
            if (!(tension[0] && tension[1] && tension[2])) {
                float add = 0;
                for (int i = 0; i < 3; i++) if (tension[i]) add += intStrRet[i];
                float value = pow(1.0f + add, -mat->nu) - 1.0f;
                for (int i = 0; i < 3; i++) if (!tension[i]) intStrRet[i] = value;
            }
        }
        return intStrRet;
    }
    Vf poissonsStrainV() {
        if (poissonsStrainInvalid) { pStrain = strainV(true); poissonsStrainInvalid = false; }
        return pStrain;
    }
    float transverseStrainSum(int ax) {
        if (mat->nu == 0) return 0;
        Vf ps = poissonsStrainV();
        switch (ax) { case X_AXIS: return ps.y + ps.z; case Y_AXIS: return ps.x + ps.z; case Z_AXIS: return ps.x + ps.y; default: return 0.0f; }
    }
    float transverseArea(int ax) {
        float size = (float)mat->nomSize;
        if (mat->nu == 0) return size * size;
        Vd ps = poissonsStrainV();
        switch (ax) {
        case X_AXIS: return (float)(size * size * (1 + ps.y) * (1 + ps.z));
        case Y_AXIS: return (float)(size * size * (1 + ps.x) * (1 + ps.z));
        case Z_AXIS: return (float)(size * size * (1 + ps.x) * (1 + ps.y));
        default: return size * size;
        }
    }
    Vd force();
    Vd moment();
    void floorForce(float dt, Vd* pTotalForce);
    void timeStep(float dt);
    Vf cornerOffset(int corner) const;
};

struct Collision {
    Voxel *pV1, *pV2;
    float penetrationStiff, dampingC;
    Vf force;
    static constexpr float envelopeRadius = 0.625f;   // VX_Collision.cpp:15
    Collision(Voxel* v1, Voxel* v2) {
        pV1 = v1; pV2 = v2;
        penetrationStiff = 2.0f / (1.0f / v1->mat->penetrationStiffness() + 1.0f / pV2->mat->penetrationStiffness());
        dampingC = 0.5f * (v1->mat->collisionDampingTranslateC() + v2->mat->collisionDampingTranslateC());
    }
    Vf contactForce(Voxel* p) const { if (p == pV1) return force; else if (p == pV2) return -force; return Vf(0, 0, 0); }
    // VX_Collision.cpp:41-56
    void updateContactForce() {
        Vf offset = (Vf)(pV2->pos - pV1->pos);
        float NomDist = (float)((pV1->baseSizeAverage() + pV2->baseSizeAverage()) * envelopeRadius);
        float RelDist = NomDist - offset.Length();
        if (RelDist > 0) {
            Vf unit = offset.Normalized();
            float relativeVelocity = pV1->velocity().Dot((Vd)unit) - pV2->velocity().Dot((Vd)unit);
            force = unit * (penetrationStiff * RelDist + dampingC * relativeVelocity);
        } else force = Vf(0, 0, 0);
    }
};

/* ---- link implementation (VX_Link.cpp) ---- */
void Link::init(Voxel* voxel1, Voxel* voxel2, MaterialLink* material) {
    bool reverseOrder = false;
    if (voxel1->ix == voxel2->ix && voxel1->iy == voxel2->iy) {
        if (voxel1->iz == voxel2->iz + 1) reverseOrder = true;
        else if (voxel1->iz + 1 == voxel2->iz) reverseOrder = false;
        else return;
        axis = Z_AXIS;
    } else if (voxel1->ix == voxel2->ix && voxel1->iz == voxel2->iz) {
        if (voxel1->iy == voxel2->iy + 1) reverseOrder = true;
        else if (voxel1->iy + 1 == voxel2->iy) reverseOrder = false;
        else return;
        axis = Y_AXIS;
    } else if (voxel1->iy == voxel2->iy && voxel1->iz == voxel2->iz) {
        if (voxel1->ix == voxel2->ix + 1) reverseOrder = true;
        else if (voxel1->ix + 1 == voxel2->ix) reverseOrder = false;
        else return;
        axis = X_AXIS;
    } else return;
    if (reverseOrder) { pVNeg = voxel2; pVPos = voxel1; } else { pVNeg = voxel1; pVPos = voxel2; }
    mat = material;
    boolStates = 0;
    valid = true;
    reset();
}
void Link::reset() {
    pos2 = angle1v = angle2v = Vd();
    angle1 = angle2 = Qd();
    forceNeg = forcePos = momentNeg = momentPos = Vd();
    strain = maxStrain = strainOffset = _stress = 0.0f;
    strainRatio = pVPos->mat->E / pVNeg->mat->E;
    smallAngle = true;
    setBoolState(LOCAL_VELOCITY_VALID, false);
    updateRestLength();
    updateTransverseInfo();
}
void Link::updateRestLength() { currentRestLength = 0.5 * (pVNeg->baseSize(axis) + pVPos->baseSize(axis)); }
void Link::updateTransverseInfo() {
    currentTransverseArea = 0.5f * (pVNeg->transverseArea(axis) + pVPos->transverseArea(axis));
    currentTransverseStrainSum = 0.5f * (pVNeg->transverseStrainSum(axis) + pVPos->transverseStrainSum(axis));
}
// VX_Link.cpp:72-113
Qd Link::orientLink() {
    pos2 = toAxisX(Vd(pVPos->pos - pVNeg->pos));
    angle1 = toAxisX(pVNeg->orient);
    angle2 = toAxisX(pVPos->orient);
    Qd totalRot = angle1.Conjugate();
    pos2 = totalRot.RotateVec3D(pos2);
    angle2 = totalRot * angle2;
    angle1 = Qd();
    float SmallTurn = (float)((std::abs(pos2.z) + std::abs(pos2.y)) / pos2.x);
    float ExtendPerc = (float)(std::abs(1 - pos2.x / currentRestLength));
    if (!smallAngle && SmallTurn < SA_BOND_BEND_RAD && ExtendPerc < SA_BOND_EXT_PERC) {
        smallAngle = true;
        setBoolState(LOCAL_VELOCITY_VALID, false);
    } else if (smallAngle && (SmallTurn > HYSTERESIS_FACTOR * SA_BOND_BEND_RAD || ExtendPerc > HYSTERESIS_FACTOR * SA_BOND_EXT_PERC)) {
        smallAngle = false;
        setBoolState(LOCAL_VELOCITY_VALID, false);
    }
    if (smallAngle) {
        pos2.x -= currentRestLength;
    } else {
        angle1.FromAngleToPosX(pos2);
        totalRot = angle1 * totalRot;
        angle2 = angle1 * angle2;
        pos2 = Vd(pos2.Length() - currentRestLength, 0, 0);
    }
    angle1v = angle1.ToRotationVector();
    angle2v = angle2.ToRotationVector();
    return totalRot;
}
// VX_Link.cpp:143-209
void Link::updateForces() {
    Vd oldPos2 = pos2, oldAngle1v = angle1v, oldAngle2v = angle2v;
    orientLink();
    Vd dPos2 = 0.5 * (pos2 - oldPos2);
    Vd dAngle1 = 0.5 * (angle1v - oldAngle1v);
    Vd dAngle2 = 0.5 * (angle2v - oldAngle2v);
    if (!mat->isXyzIndependent() || currentTransverseStrainSum != 0) updateTransverseInfo();
    _stress = updateStrain((float)(pos2.x / currentRestLength));
    if (isFailed()) { forceNeg = forcePos = momentNeg = momentPos = Vd(0, 0, 0); return; }
    float b1 = mat->_b1, b2 = mat->_b2, b3 = mat->_b3, a2 = mat->_a2;
    forceNeg = Vd(_stress * currentTransverseArea, b1 * pos2.y - b2 * (angle1v.z + angle2v.z), b1 * pos2.z + b2 * (angle1v.y + angle2v.y));
    forcePos = -forceNeg;
    momentNeg = Vd(a2 * (angle2v.x - angle1v.x), -b2 * pos2.z - b3 * (2 * angle1v.y + angle2v.y), b2 * pos2.y - b3 * (2 * angle1v.z + angle2v.z));
    momentPos = Vd(a2 * (angle1v.x - angle2v.x), -b2 * pos2.z - b3 * (angle1v.y + 2 * angle2v.y), b2 * pos2.y - b3 * (angle1v.z + 2 * angle2v.z));
    if (isLocalVelocityValid()) {
        float sqA1 = mat->_sqA1, sqA2xIp = mat->_sqA2xIp, sqB1 = mat->_sqB1, sqB2xFMp = mat->_sqB2xFMp, sqB3xIp = mat->_sqB3xIp;
        Vd posCalc(sqA1 * dPos2.x, sqB1 * dPos2.y - sqB2xFMp * (dAngle1.z + dAngle2.z), sqB1 * dPos2.z + sqB2xFMp * (dAngle1.y + dAngle2.y));
        forceNeg += pVNeg->dampingMultiplier() * posCalc;
        forcePos -= pVPos->dampingMultiplier() * posCalc;
        momentNeg -= 0.5 * pVNeg->dampingMultiplier() *
                     Vd(-sqA2xIp * (dAngle2.x - dAngle1.x), sqB2xFMp * dPos2.z + sqB3xIp * (2 * dAngle1.y + dAngle2.y),
                        -sqB2xFMp * dPos2.y + sqB3xIp * (2 * dAngle1.z + dAngle2.z));
        momentPos -= 0.5 * pVPos->dampingMultiplier() *
                     Vd(sqA2xIp * (dAngle2.x - dAngle1.x), sqB2xFMp * dPos2.z + sqB3xIp * (dAngle1.y + 2 * dAngle2.y),
                        -sqB2xFMp * dPos2.y + sqB3xIp * (dAngle1.z + 2 * dAngle2.z));
    } else setBoolState(LOCAL_VELOCITY_VALID, true);
    if (!smallAngle) {
        forceNeg = angle1.RotateVec3DInv(forceNeg);
        momentNeg = angle1.RotateVec3DInv(momentNeg);
    }
    forcePos = angle2.RotateVec3DInv(forcePos);
    momentPos = angle2.RotateVec3DInv(momentPos);
    toAxisOriginal(&forceNeg);
    toAxisOriginal(&forcePos);
    toAxisOriginal(&momentNeg);
    toAxisOriginal(&momentPos);
}
// VX_Link.cpp:212-241
float Link::updateStrain(float axialStrain) {
    strain = axialStrain;
    if (mat->linear) {
        if (axialStrain > maxStrain) maxStrain = axialStrain;
        return mat->stress(axialStrain, currentTransverseStrainSum);
    } else {
        float returnStress;
        if (axialStrain > maxStrain) {
            maxStrain = axialStrain;
            returnStress = mat->stress(axialStrain, currentTransverseStrainSum);
            if (mat->nu != 0.0f) strainOffset = maxStrain - mat->stress(axialStrain) / (mat->_eHat * (1 - mat->nu));
            else strainOffset = maxStrain - returnStress / mat->E;
        } else {
            float relativeStrain = axialStrain - strainOffset;
            if (mat->nu != 0.0f) returnStress = mat->stress(relativeStrain, currentTransverseStrainSum, true);
            else returnStress = mat->E * relativeStrain;
        }
        return returnStress;
    }
}
// VX_Link.cpp:251-259
float Link::axialStiffness() {
    if (mat->isXyzIndependent()) return mat->_a1;
    updateRestLength();
    updateTransverseInfo();
    return (float)(mat->_eHat * currentTransverseArea / ((strain + 1) * currentRestLength));
}

/* ---- voxel implementation (VX_Voxel.cpp) ---- */
Vd Voxel::force() {
    Vd totalForce(0, 0, 0);
    for (int i = 0; i < 6; i++) if (links[i]) totalForce += links[i]->force(isNegative(i));
    totalForce = orient.RotateVec3D(totalForce);
    if (ext) totalForce += ext->extForce;
    totalForce -= velocity() * mat->globalDampingTranslateC();
    totalForce.z += mat->gravityForce();
    if (isCollisionsEnabled()) {
        for (Collision* c : *colWatch) totalForce -= c->contactForce(this);
    }
    if (extras) totalForce += ciliaForce;
    return totalForce;
}
Vd Voxel::moment() {
    Vd totalMoment(0, 0, 0);
    for (int i = 0; i < 6; i++) if (links[i]) totalMoment += links[i]->moment(isNegative(i));
    totalMoment = orient.RotateVec3D(totalMoment);
    if (ext) totalMoment += ext->extMoment;
    totalMoment -= angularVelocity() * mat->globalDampingRotateC();
    return totalMoment;
}
// VX_Voxel.cpp:272-297
void Voxel::floorForce(float dt, Vd* pTotalForce) {
    (void)dt;
    float CurPenetration = floorPenetration();
    if (CurPenetration >= 0) {
        Vd vel = velocity();
        Vd horizontalVel(vel.x, vel.y, 0);
        float normalForce = mat->penetrationStiffness() * CurPenetration;
        pTotalForce->z += normalForce - mat->collisionDampingTranslateC() * vel.z;
        if (isFloorStaticFriction()) {
            float surfaceForceSq = (float)(pTotalForce->x * pTotalForce->x + pTotalForce->y * pTotalForce->y);
            float frictionForceSq = (mat->muStatic * normalForce) * (mat->muStatic * normalForce);
            if (surfaceForceSq > frictionForceSq) setFloorStaticFriction(false);
        } else {
            *pTotalForce -= mat->muKinetic * normalForce * horizontalVel.Normalized();
        }
    } else setFloorStaticFriction(false);
}
// VX_Voxel.cpp:163-230
void Voxel::timeStep(float dt) {
    previousDt = dt;
    if (dt == 0.0f) return;
    if (ext && ext->isFixedAll()) {
        pos = originalPosition() + ext->extTranslation;
        orient = ext->rotationQuat();
        haltMotion();
        return;
    }
    Vd curForce = force();
    if (extras) {
        curForce.x += fieldForce.x;
        curForce.y += fieldForce.y;
        curForce.z += fieldForce.z;
    }
    Vd fricForce = curForce;
    if (isFloorEnabled()) floorForce(dt, &curForce);
    fricForce = curForce - fricForce;
    linMom += curForce * dt;
    Vd translate(linMom * (dt * mat->_massInverse));
    if (isFloorEnabled() && floorPenetration() >= 0) {
        double work = fricForce.x * translate.x + fricForce.y * translate.y;
        double hKe = 0.5 * mat->_massInverse * (linMom.x * linMom.x + linMom.y * linMom.y);
        if (hKe + work <= 0) setFloorStaticFriction(true);
        if (isFloorStaticFriction()) {
            linMom.x = linMom.y = 0;
            translate.x = translate.y = 0;
        }
    } else setFloorStaticFriction(false);
    pos += translate;
    Vd curMoment = moment();
    angMom += curMoment * dt;
    orient = Qd(angMom * (dt * mat->_momentInertiaInverse)) * orient;
    if (ext) {
        double size = mat->nomSize;
        if (ext->isFixed(X_TRANSLATE)) { pos.x = ix * size + ext->extTranslation.x; linMom.x = 0; }
        if (ext->isFixed(Y_TRANSLATE)) { pos.y = iy * size + ext->extTranslation.y; linMom.y = 0; }
        if (ext->isFixed(Z_TRANSLATE)) { pos.z = iz * size + ext->extTranslation.z; linMom.z = 0; }
        if (ext->isFixedAnyRotation()) {
            if (ext->isFixedAllRotation()) {
                orient = ext->rotationQuat();
                angMom = Vd();
            } else {
                Vd tmpRotVec = orient.ToRotationVector();
                if (ext->isFixed(X_ROTATE)) { tmpRotVec.x = 0; angMom.x = 0; }
                if (ext->isFixed(Y_ROTATE)) { tmpRotVec.y = 0; angMom.y = 0; }
                if (ext->isFixed(Z_ROTATE)) { tmpRotVec.z = 0; angMom.z = 0; }
                orient.FromRotationVector(tmpRotVec);
            }
        }
    }
    poissonsStrainInvalid = true;
}
// VX_Voxel.cpp:147-160
Vf Voxel::cornerOffset(int corner) const {
    Vd strains;
    for (int i = 0; i < 3; i++) {
        bool posLink = corner & (1 << (2 - i)) ? true : false;
        Link* pL = links[2 * i + (posLink ? 0 : 1)];
        if (pL && !pL->isFailed()) strains[i] = (1 + pL->axialStrain(posLink)) * (posLink ? 1 : -1);
        else strains[i] = posLink ? 1.0 : -1.0;
    }
    return (Vf)((0.5 * baseSize()).Scale(strains));
}

/* ------------------------------------------------------------------ engine (Voxelyze.cpp) */
typedef std::tuple<int, int, int> Key;

struct Engine {
    double voxSize = 0.001;
    float currentTime = 0, ambientTemp = 0, grav = 0;
    bool floor = false, collisions = false;
    bool detach = false, surfaceChanged = false;   // VX3 detach of failed links
    bool attach = false;                           // VX3 attach of touching sticky voxels
    int safetyGuard = 500;
    std::function<void()> afterContacts;           // VX3 target contacts, between the contact and voxel updates
    float boundingRadius = 0.75f, watchDistance = 1.0f;
    std::vector<MaterialVoxel*> voxelMats;
    std::list<MaterialLink*> linkMats;
    std::map<Key, Voxel*> voxels;
    std::vector<Voxel*> voxelsList;
    std::map<Key, Link*> links[3];
    std::vector<Link*> linksList;
    std::vector<Collision*> collisionsList;
    bool collisionsStale = true, nearbyStale = true;
    int colInterval = 0;        // > 0: also regenerate at every step index that is a multiple of it
    long colStepIndex = 0;      // zero-based index of the step being run (set by the VX3 driver)

    ~Engine() {
        for (Collision* c : collisionsList) delete c;
        for (Link* l : linksList) delete l;
        for (Voxel* v : voxelsList) delete v;
        for (MaterialVoxel* m : voxelMats) delete m;
        for (MaterialLink* m : linkMats) delete m;
    }
    Voxel* voxelAt(int x, int y, int z) const { auto it = voxels.find(Key(x, y, z)); return it == voxels.end() ? nullptr : it->second; }
    static int xLO(int d) { return d == X_NEG ? -1 : 0; }
    static int yLO(int d) { return d == Y_NEG ? -1 : 0; }
    static int zLO(int d) { return d == Z_NEG ? -1 : 0; }
    static int xVO(int d) { return d == X_NEG ? -1 : (d == X_POS ? 1 : 0); }
    static int yVO(int d) { return d == Y_NEG ? -1 : (d == Y_POS ? 1 : 0); }
    static int zVO(int d) { return d == Z_NEG ? -1 : (d == Z_POS ? 1 : 0); }
    Link* linkAt(int x, int y, int z, int d) const {
        auto& m = links[Voxel::toAxis(d)];
        auto it = m.find(Key(x + xLO(d), y + yLO(d), z + zLO(d)));
        return it == m.end() ? nullptr : it->second;
    }
    MaterialVoxel* addMaterial() {
        MaterialVoxel* p = new MaterialVoxel(1e6f, 1e3f, voxSize);
        p->gravMult = grav;
        voxelMats.push_back(p);
        return p;
    }
    // Voxelyze.cpp:626-641
    MaterialLink* combinedMaterial(MaterialVoxel* m1, MaterialVoxel* m2) {
        for (MaterialLink* t : linkMats)
            if ((t->vox1Mat == m1 && t->vox2Mat == m2) || (t->vox1Mat == m2 && t->vox2Mat == m1)) return t;
        MaterialLink* n = new MaterialLink(m1, m2);
        linkMats.push_back(n);
        m1->dependentMaterials.push_back(n);
        if (m1 != m2) m2->dependentMaterials.push_back(n);
        return n;
    }
    // Voxelyze.cpp:508-539
    Link* addLink(int x, int y, int z, int d) {
        Link* pL = linkAt(x, y, z, d);
        if (pL) return pL;
        Voxel* v1 = voxelAt(x, y, z);
        Voxel* v2 = voxelAt(x + xVO(d), y + yVO(d), z + zVO(d));
        if (!v1 || !v2) return nullptr;
        MaterialLink* m = combinedMaterial(v1->mat, v2->mat);
        pL = new Link();
        pL->init(v1, v2, m);
        linksList.push_back(pL);
        links[Voxel::toAxis(d)][Key(x + xLO(d), y + yLO(d), z + zLO(d))] = pL;
        v1->links[d] = pL; v1->updateSurface();
        v2->links[Voxel::toOpposite(d)] = pL; v2->updateSurface();
        return pL;
    }
    void removeLink(int x, int y, int z, int d) {
        Link* pL = linkAt(x, y, z, d);
        if (!pL) return;
        links[Voxel::toAxis(d)].erase(Key(x + xLO(d), y + yLO(d), z + zLO(d)));
        for (auto it = linksList.begin(); it != linksList.end(); ++it) if (*it == pL) { linksList.erase(it); break; }
        Voxel* v1 = voxelAt(x, y, z);
        if (v1) { v1->links[d] = nullptr; v1->updateSurface(); }
        Voxel* v2 = voxelAt(x + xVO(d), y + yVO(d), z + zVO(d));
        if (v2) { v2->links[Voxel::toOpposite(d)] = nullptr; v2->updateSurface(); }
        delete pL;
    }
    // Voxelyze.cpp:439-461
    Voxel* addVoxel(MaterialVoxel* m, int x, int y, int z) {
        nearbyStale = collisionsStale = true;
        Voxel* pV = new Voxel();
        pV->mat = m; pV->ix = x; pV->iy = y; pV->iz = z;
        pV->reset();
        voxels[Key(x, y, z)] = pV;
        pV->index = (int)voxelsList.size();
        voxelsList.push_back(pV);
        pV->pos = Vd(x * voxSize, y * voxSize, z * voxSize);
        pV->enableFloor(floor);
        pV->setTemperature(ambientTemp);
        pV->enableCollisions(collisions);
        for (int i = 0; i < 6; i++) addLink(x, y, z, i);
        return pV;
    }
    // Voxelyze.cpp:422-437,485-498 + CVX_Voxel::replaceMaterial (VX_Voxel.cpp:79-91)
    Voxel* setVoxel(MaterialVoxel* m, int x, int y, int z) {
        Voxel* pV = voxelAt(x, y, z);
        if (pV) {
            collisionsStale = true;
            // CVX_Voxel::replaceMaterial
            pV->linMom = pV->linMom * (double)(m->_mass / pV->mat->_mass);
            pV->angMom = pV->angMom * (double)(m->_momentInertia / pV->mat->_momentInertia);
            pV->setFloorStaticFriction(false);
            pV->poissonsStrainInvalid = true;
            pV->mat = m;
            for (int i = 0; i < 6; i++) { removeLink(x, y, z, i); addLink(x, y, z, i); }
            return pV;
        }
        return addVoxel(m, x, y, z);
    }
    void setVoxelSize(double s) {
        double scaleFactor = s / voxSize;
        voxSize = s;
        for (MaterialVoxel* m : voxelMats) m->setNominalSize(s);
        for (Voxel* v : voxelsList) { v->pos = v->pos * scaleFactor; v->haltMotion(); v->setFloorStaticFriction(false); }
        for (Link* l : linksList) l->reset();
        collisionsStale = true;
    }
    void setGravity(float g) { grav = g; for (MaterialVoxel* m : voxelMats) m->gravMult = grav; }
    void enableFloor(bool e) { floor = e; for (Voxel* v : voxelsList) v->enableFloor(e); }
    void clearCollisions() {
        for (Collision* c : collisionsList) delete c;
        collisionsList.clear();
        for (Voxel* v : voxelsList) if (v->colWatch) v->colWatch->clear();
    }
    void enableCollisions(bool e) {
        if (collisions == e) return;
        collisions = e;
        for (Voxel* v : voxelsList) v->enableCollisions(e);
        if (!collisions) clearCollisions();
        collisionsStale = true;
    }
    void setAmbientTemperature(float t, bool all) {
        ambientTemp = t;
        if (all) for (Voxel* v : voxelsList) v->setTemperature(t);
    }
    // Voxelyze.cpp:286-311
    float recommendedTimeStep() {
        float MaxFreq2 = 0.0f;
        for (Link* pL : linksList) {
            float m1 = pL->pVNeg->mat->_mass, m2 = pL->pVPos->mat->_mass;
            float thisMaxFreq2 = pL->axialStiffness() / (m1 < m2 ? m1 : m2);
            if (thisMaxFreq2 > MaxFreq2) MaxFreq2 = thisMaxFreq2;
        }
        if (MaxFreq2 <= 0.0f) {
            for (Voxel* v : voxelsList) {
                float t = v->mat->E * v->mat->nomSize / v->mat->_mass;
                if (t > MaxFreq2) MaxFreq2 = t;
            }
        }
        if (MaxFreq2 <= 0.0f) return 0.0f;
        return 1.0f / (6.283185f * sqrt(MaxFreq2));
    }
    void resetTime() {
        currentTime = 0.0f;
        collisionsStale = true;
        nearbyStale = true;
        for (Voxel* v : voxelsList) v->reset();
        for (Link* l : linksList) l->reset();
    }
    // Voxelyze.cpp:394-417.  The breadth-first order and the resulting set are the
    // reference's; membership is tracked with a per-voxel stamp instead of a linear
    // search of the list (same result, linear time).
    std::vector<int> nearStamp;
    int nearTick = 0;
    void generateNearby(Voxel* self, int linkDepth) {
        if (nearStamp.size() != voxelsList.size()) { nearStamp.assign(voxelsList.size(), 0); nearTick = 0; }
        const int tick = ++nearTick;
        std::vector<Voxel*> allNearby;
        allNearby.push_back(self);
        nearStamp[self->index] = tick;
        int iCurrent = 0;
        for (int k = 0; k < linkDepth; k++) {
            int iPassEnd = (int)allNearby.size();
            while (iCurrent != iPassEnd) {
                Voxel* pV = allNearby[iCurrent++];
                for (int i = 0; i < 6; i++) {
                    Voxel* pV2 = pV->adjacentVoxel(i);
                    if (pV2 && nearStamp[pV2->index] != tick) { nearStamp[pV2->index] = tick; allNearby.push_back(pV2); }
                }
            }
        }
        self->nearby->clear();
        for (Voxel* pV : allNearby) if (pV->isSurface() && pV != self) self->nearby->push_back(pV);
    }
    // Voxelyze.cpp:726-751.  The reference tests every later voxel b of every surface
    // voxel a; here the candidates b come from a uniform grid of cell edge slightly
    // above the threshold radius (every pair within the radius lies in neighbouring
    // cells, whatever the rounding of pos/cell), and each a takes its partners in
    // ascending list order -- the same pairs, in the same order, on collisionsList and
    // on every voxel's colWatch list.
    void regenerateCollisions(float threshRadiusSq) {
        clearCollisions();
        const double cell = std::sqrt((double)threshRadiusSq) * 1.0001 + 1e-300;
        std::map<std::tuple<long, long, long>, std::vector<int>> grid;
        auto key = [&](const Vd& p) {
            return std::make_tuple((long)std::floor(p.x / cell), (long)std::floor(p.y / cell), (long)std::floor(p.z / cell));
        };
        for (size_t a = 0; a < voxelsList.size(); a++)
            if (!voxelsList[a]->isInterior()) grid[key(voxelsList[a]->pos)].push_back((int)a);
        if (nearStamp.size() != voxelsList.size()) { nearStamp.assign(voxelsList.size(), 0); nearTick = 0; }
        std::vector<int> cand;
        for (size_t a = 0; a < voxelsList.size(); a++) {
            Voxel* pV1 = voxelsList[a];
            if (pV1->isInterior()) continue;
            *pV1->lastColWatchPosition = (Vf)pV1->pos;
            const int tick = ++nearTick;
            for (Voxel* n : *pV1->nearby) nearStamp[n->index] = tick;
            cand.clear();
            const auto c = key(pV1->pos);
            for (long dz = -1; dz <= 1; dz++)
                for (long dy = -1; dy <= 1; dy++)
                    for (long dx = -1; dx <= 1; dx++) {
                        auto it = grid.find(std::make_tuple(std::get<0>(c) + dx, std::get<1>(c) + dy, std::get<2>(c) + dz));
                        if (it == grid.end()) continue;
                        for (int b : it->second) if (b > (int)a) cand.push_back(b);
                    }
            std::sort(cand.begin(), cand.end());
            for (int b : cand) {
                Voxel* pV2 = voxelsList[b];
                if ((pV1->pos - pV2->pos).Length2() > threshRadiusSq || nearStamp[pV2->index] == tick) continue;
                Collision* col = new Collision(pV1, pV2);
                collisionsList.push_back(col);
                pV1->colWatch->push_back(col);
                pV2->colWatch->push_back(col);
            }
        }
        collisionsStale = false;
    }
    // Voxelyze.cpp:671-711
    void updateCollisions() {
        float watchRadiusVx = 2 * boundingRadius + watchDistance;
        float watchRadiusMm = (float)(voxSize * watchRadiusVx);
        float recalcDist = (float)(voxSize * watchDistance / 2);
        if (nearbyStale) {
            for (Voxel* v : voxelsList) generateNearby(v, (int)(watchRadiusVx * 2));
            nearbyStale = false;
            collisionsStale = true;
        }
        for (Voxel* pV : voxelsList) {
            if (pV->isSurface() && (pV->pos - *pV->lastColWatchPosition).Length2() > recalcDist * recalcDist) collisionsStale = true;
        }
        if (colInterval > 0 && colStepIndex % colInterval == 0) collisionsStale = true;
        if (collisionsStale) regenerateCollisions(watchRadiusMm * watchRadiusMm);
        for (Collision* c : collisionsList) c->updateContactForce();
    }
    // is_neighbor (VX3_VoxelyzeKernel.cu:651-680)
    static bool linkedWithin(Voxel* a, Voxel* t, Link* in, int depth) {
        if (a == t) return true;
        if (depth <= 0) return false;
        for (int i = 0; i < 6; i++) {
            Link* l = a->links[i];
            if (!l || l == in) continue;
            if (linkedWithin(l->pVNeg == a ? l->pVPos : l->pVNeg, t, l, depth - 1)) return true;
        }
        return false;
    }
    // the attach half of handle_collision_attachment (VX3_VoxelyzeKernel.cu:724-826) over the
    // pairs of gpu_update_attach (:833-843), surface voxels first > second, taken one after
    // another in that order
    void updateAttach() {
        std::vector<Voxel*> surf;
        for (Voxel* v : voxelsList)
            if (v->isSurface() && !v->removed) surf.push_back(v);
        for (size_t i = 1; i < surf.size(); i++)
            for (size_t j = 0; j < i; j++) {
                Voxel *v1 = surf[i], *v2 = surf[j];
                if (v1->mat->fixed && v2->mat->fixed) continue;
                const double wd = (v1->baseSizeAverage() + v2->baseSizeAverage()) * 0.625;
                const Vd diff = v1->pos - v2->pos;
                if (diff.x > wd || diff.x < -wd || diff.y > wd || diff.y < -wd || diff.z > wd || diff.z < -wd) continue;
                if (diff.Length() > wd) continue;
                if (linkedWithin(v1, v2, nullptr, 1)) continue;
                if (!v1->enableAttach || !v2->enableAttach) continue;
                if (v1->mat->fixed || v2->mat->fixed) continue;
                if (v1->mat != v2->mat || !v1->mat->sticky) continue;
                if (linkedWithin(v1, v2, nullptr, 5)) continue;
                const Vd e = v1->pos - v2->pos;
                const Vd ea = v1->orient.RotateVec3DInv(-e);
                const Vd f(fabs(ea.x), fabs(ea.y), fabs(ea.z));
                int ax;
                double c;
                if (f.x >= f.y && f.x >= f.z) { ax = X_AXIS; c = ea.x; }
                else if (f.y >= f.x && f.y >= f.z) { ax = Y_AXIS; c = ea.y; }
                else { ax = Z_AXIS; c = ea.z; }
                const bool rev = c < 0;
                const int d1 = 2 * ax + (rev ? 1 : 0), d2 = d1 ^ 1;
                if (v1->links[d1] || v2->links[d2]) continue;
                Link* L = new Link();
                L->axis = ax;
                L->pVPos = rev ? v1 : v2;   // VX3_Link device ctor (VX3_Link.cu:30-55)
                L->pVNeg = rev ? v2 : v1;
                L->mat = combinedMaterial(v1->mat, v2->mat);
                L->boolStates = 0;
                L->valid = true;
                L->reset();
                L->fresh = true;
                L->isNewLink = safetyGuard;
                v1->links[d1] = L;
                v2->links[d2] = L;
                linksList.push_back(L);
                surfaceChanged = true;
            }
    }
    // Voxelyze.cpp:251-284 (fixed-material skips follow VX3_VoxelyzeKernel.cu:572,588)
    bool doTimeStep(float dt) {
        if (dt == 0) return true;
        else if (dt < 0) dt = recommendedTimeStep();
        bool Diverged = false;
        // VX3 runs updateAttach after the link pass (VX3_VoxelyzeKernel.cu:289-291); a link it
        // makes carries no force until the next step's link pass, so attaching before this
        // step's link pass and skipping the new links there is the same
        if (attach) updateAttach();
        for (Link* l : linksList) {
            if (l->pVPos->mat->fixed && l->pVNeg->mat->fixed) continue;
            if (l->gone) continue;
            if (l->fresh) { l->fresh = false; continue; }
            l->updateForces();
            if (l->isNewLink) {   // VX3_Link.cu:206-213
                l->forceNeg = l->forceNeg * 0.01; l->forcePos = l->forcePos * 0.01;
                l->momentNeg = l->momentNeg * 0.01; l->momentPos = l->momentPos * 0.01;
                l->isNewLink -= 1;
            }
            if (l->axialStrain() > 100) Diverged = true;
        }
        if (Diverged) return false;
        // gpu_update_detach (VX3_VoxelyzeKernel.cu:946-967); the voxels' surface state is
        // refreshed at the next step (regenerateSurfaceVoxels, :281-285)
        if (detach)
            for (Link* l : linksList) {
                if (l->gone || !l->isFailed()) continue;
                l->gone = true;
                for (int i = 0; i < 6; i++) {
                    if (l->pVNeg->links[i] == l) l->pVNeg->links[i] = nullptr;
                    if (l->pVPos->links[i] == l) l->pVPos->links[i] = nullptr;
                }
                surfaceChanged = true;
            }
        if (collisions) updateCollisions();
        if (afterContacts) afterContacts();
        for (Voxel* v : voxelsList) {
            if (v->mat->fixed) continue;
            if (v->removed) continue;   // VX3_VoxelyzeKernel.cu:586
            v->timeStep(dt);
        }
        currentTime += dt;
        return true;
    }
};

/* ------------------------------------------------------------------ simulation (import + VX3 driver) */
struct Sim {
    Engine vx;
    std::vector<int> matIndexOfVoxelMat;   // voxel material list order == desc order
    // VX3 driver state
    double currentTime = 0.0, OptimalDt = 0.0, DtFrac = 0.9, dtFixed = 0.0;
    long CurStepCount = 0;
    bool vary = false, enableExpansion = false;
    double TempAmplitude = 0, TempPeriod = 0;
    Vd initialCoM, currentCoM, comHist[2];
    int angleSampleTimes = 0;
    double recentAngle = 0;
    int state = VX3H_SIM_RUNNING;
    int maxSteps = 0;
    std::vector<vx3h_token> stopTok, fitTok;
    float lastDt = 0;
    VoxcadStop voxcad;   // VoxCAD stop-condition family and equilibrium mode
    int hits = 0;        // collisionCount (the MathTree variable hit)
    // computeTargetCloseness (VX3_VoxelyzeKernel.cu:545-562) over the voxels of isTarget
    // materials in voxel order (registerTargets, :536-543), at every center-of-mass sample
    double closeR = 0, targetCloseness = 0;
    int numClosePairs = 0;
    std::vector<Voxel*> closeTargets;
    void computeTargetCloseness() {
        if (closeR == 0) return;
        double ret = 0;
        numClosePairs = 0;
        for (size_t i = 0; i < closeTargets.size(); i++)
            for (size_t j = i + 1; j < closeTargets.size(); j++) {
                const double distance = closeTargets[i]->pos.Dist(closeTargets[j]->pos);
                if (distance < closeR) numClosePairs++;
                ret += 1 / distance;
            }
        targetCloseness = ret;
    }
    std::vector<vx3h_token> attachTok[5];   // VX3 attach conditions

    // handle_collision_attachment (VX3_VoxelyzeKernel.cu:682-722) over the watched pairs:
    // a target and another voxel, not both fixed, with centres within 0.625 x the sum
    // of their average base sizes, count a hit; with signals the other voxel receives 100
    void targetContacts() {
        for (Collision* c : vx.collisionsList) {
            Voxel *a = c->pV1, *b = c->pV2;
            if (a->target == b->target) continue;
            if (a->mat->fixed && b->mat->fixed) continue;
            const double wd = (a->baseSizeAverage() + b->baseSizeAverage()) * 0.625;
            const Vd d = a->pos - b->pos;
            if (d.x > wd || d.x < -wd || d.y > wd || d.y < -wd || d.z > wd || d.z < -wd) continue;
            if (d.Length() > wd) continue;
            hits++;
            if (signals) (a->target ? b : a)->receiveSignal(100, currentTime, true);
        }
    }
    // VX3 secondary experiment (VX3_VoxelyzeKernel.cu:336-350)
    bool secondary = false, reinitDone = false;
    double reinitAfter = 0;
    std::vector<double> matRemoveAfter;
    std::vector<bool> matRemoved;
    std::vector<Vd> initPos;   // d_initialPosition: lattice positions, or those saved at the reset

    // removeVoxels (VX3_VoxelyzeKernel.cu:365-397)
    void removeVoxels() {
        for (size_t i = 0; i < matRemoveAfter.size(); i++) {
            if (matRemoved[i] || !(matRemoveAfter[i] > 0) || !(matRemoveAfter[i] < currentTime)) continue;
            for (Voxel* v : vx.voxelsList) {
                if (v->mat != vx.voxelMats[i] || v->removed) continue;
                v->removed = true;
                for (int k = 0; k < 6; k++) {
                    Link* l = v->links[k];
                    if (!l) continue;
                    l->gone = true;
                    Voxel* nb = l->pVNeg == v ? l->pVPos : l->pVNeg;
                    for (int m = 0; m < 6; m++)
                        if (nb->links[m] == l) { nb->links[m] = nullptr; break; }
                    v->links[k] = nullptr;
                }
            }
            matRemoved[i] = true;
            vx.surfaceChanged = true;
        }
    }
    // VX3 cilia and force field
    bool extras = false, ciliaOn = false, signals = false;
    std::vector<double> matCilia, matCiliaOn;   // per material (voxelMats order)
    std::vector<vx3h_token> ffTok[3];

    // gpu_update_cilia_force (VX3_VoxelyzeKernel.cu:846-859) with localSignal 0, and the
    // force field of VX3_Voxel::timeStep (VX3_Voxel.cu:178-184), from the state at the
    // start of the voxel update
    void updateExtras() {
        if (!extras) return;
        const int nv = (int)vx.voxelsList.size();
        for (Voxel* v : vx.voxelsList) {
            v->ciliaForce = Vd(0, 0, 0);
            const int mi = matOf(v);   // voxel materials are in description order
            if (ciliaOn && v->isSurface() && !v->removed && matCilia[mi] != 0 && !(matCiliaOn[mi] > currentTime)) {
                const double localSignal = signals ? v->localSignal : 0.0;
                v->ciliaForce = v->orient.RotateVec3D(v->baseCilia + v->shiftCilia * localSignal) * matCilia[mi];
            }
            double f[3];
            for (int k = 0; k < 3; k++)
                f[k] = ffTok[k].empty() ? 0.0
                                        : mathtree_eval(ffTok[k].data(), (int)ffTok[k].size(), v->pos.x, v->pos.y, v->pos.z, hits,
                                                        currentTime, recentAngle, targetCloseness, numClosePairs, nv);
            v->fieldForce = Vd(f[0], f[1], f[2]);
        }
    }
    int matOf(const Voxel* v) const {
        for (size_t i = 0; i < vx.voxelMats.size(); i++) if (vx.voxelMats[i] == v->mat) return (int)i;
        return 0;
    }

    // VX3_VoxelyzeKernel.cu:477-493
    void updateCurrentCenterOfMass() {
        double TotalMass = 0;
        Vd Sum(0, 0, 0);
        for (Voxel* v : vx.voxelsList) {
            if (!v->mat->isMeasured) continue;
            double ThisMass = v->mat->_mass;
            Sum += v->pos * ThisMass;
            TotalMass += ThisMass;
        }
        if (TotalMass == 0) { currentCoM = Vd(); return; }
        currentCoM = Sum / TotalMass;
    }
    // VX3_VoxelyzeKernel.cu:184-217
    double recommendedTimeStepVX3() {
        double MaxFreq2 = 0.0f;
        for (Link* pL : vx.linksList) {
            double m1 = pL->pVNeg->mat->_mass, m2 = pL->pVPos->mat->_mass;
            double thisMaxFreq2 = pL->axialStiffness() / (m1 < m2 ? m1 : m2);
            if (thisMaxFreq2 > MaxFreq2) MaxFreq2 = thisMaxFreq2;
        }
        if (MaxFreq2 <= 0.0f) {
            for (Voxel* v : vx.voxelsList) {
                double t = v->mat->E * v->mat->nomSize / v->mat->_mass;
                if (t > MaxFreq2) MaxFreq2 = t;
            }
        }
        if (MaxFreq2 <= 0.0f) return 0.0f;
        return 1.0f / (6.283185f * sqrt(MaxFreq2));
    }
    // VX3_VoxelyzeKernel.cu:219-235,625-650
    void updateTemperature() {
        if (!vary || !(TempPeriod > 0)) return;
        for (Voxel* v : vx.voxelsList) {
            if (v->removed) continue;   // VX3_VoxelyzeKernel.cu:632
            if (v->mat->thermalOnAfter > currentTime) continue;
            if (v->mat->fixed) continue;
            double T = TempAmplitude * sin(2 * 3.1415926f * (currentTime / TempPeriod + v->phaseOffset));
            if (!enableExpansion) { if (T > 0) T = 0; }
            v->setTemperature((float)T);
        }
    }
    bool stopConditionMet() {
        if (maxSteps > 0 && CurStepCount >= maxSteps) return true;
        if (!stopTok.empty() &&
            mathtree_eval(stopTok.data(), (int)stopTok.size(), currentCoM.x, currentCoM.y, currentCoM.z, hits, currentTime, recentAngle,
                          targetCloseness, numClosePairs,
                          (int)vx.voxelsList.size()) > 0)
            return true;
        return voxcad.type != 0 && voxcad.met(CurStepCount, currentTime, TempPeriod, (int)vx.voxelsList.size());
    }
    // VX3_VoxelyzeKernel.cu:237-359
    bool doTimeStepVX3() {
        updateTemperature();
        CurStepCount++;
        float dt;
        if (dtFixed > 0) dt = (float)dtFixed;
        else {
            if (!OptimalDt) OptimalDt = recommendedTimeStepVX3();
            if (OptimalDt < 1e-10) OptimalDt = 1e-10;
            dt = DtFrac * OptimalDt;
        }
        lastDt = dt;
        if (vx.surfaceChanged) {
            vx.surfaceChanged = false;
            for (Voxel* v : vx.voxelsList) v->updateSurface();
            // VX3 topology changes (attach, detach, voxel removal) of the last step: with
            // collisions, the nearby lists follow the links present now (generateNearby on
            // the current topology, then a collision regeneration), as CVoxelyze does when
            // voxels are added or removed (Voxelyze.cpp:442,466,678-684)
            if (vx.collisions) vx.nearbyStale = true;
        }
        vx.colStepIndex = CurStepCount - 1;
        if (!vx.doTimeStep(dt)) return false;   // target contacts and extra forces run inside, after the contact update
        if (vx.attach)   // VX3_VoxelyzeKernel.cu:608-621, after each voxel's update; a missing condition does not restrict
            for (Voxel* v : vx.voxelsList) {
                if (v->mat->fixed || v->removed) continue;
                bool on = true;
                for (int k = 0; k < 5 && on; k++)
                    if (!attachTok[k].empty())
                        on = mathtree_eval(attachTok[k].data(), (int)attachTok[k].size(), v->pos.x, v->pos.y, v->pos.z, hits,
                                           currentTime, recentAngle, targetCloseness, numClosePairs,
                                           (int)vx.voxelsList.size()) > 0;
                v->enableAttach = on;
            }
        if (signals) {   // every moving voxel on its own state, then every voxel takes its neighbours' posts
            for (Voxel* v : vx.voxelsList)
                if (!v->mat->fixed && !v->removed) v->signalsOwn(currentTime);
            for (Voxel* v : vx.voxelsList) {
                if (v->removed) continue;
                for (int d = 0; d < 6; d++) {
                    Link* l = v->links[d];
                    if (!l) continue;
                    Voxel* nb = l->pVNeg == v ? l->pVPos : l->pVNeg;
                    if (nb->postedValue > 0) v->receiveSignal(nb->postedValue, nb->postedTime, false);
                }
            }
        }
        int CycleStep = int(TempPeriod / dt);
        if (CycleStep > 0 && CurStepCount % CycleStep == 0) {
            angleSampleTimes++;
            comHist[0] = comHist[1];
            comHist[1] = currentCoM;
            updateCurrentCenterOfMass();
            Vd A = comHist[0], B = comHist[1], C = currentCoM;
            if (B == C || A == B || angleSampleTimes < 3) recentAngle = 0;
            else recentAngle = acos((B - A).Dot(C - B) / (B.Dist(A) * C.Dist(B)));
            computeTargetCloseness();
        }
        if (secondary) {
            removeVoxels();
            if (!reinitDone && reinitAfter < currentTime) {
                reinitDone = true;
                initialCoM = currentCoM;   // InitializeCenterOfMass
                for (size_t i = 0; i < vx.voxelsList.size(); i++) initPos[i] = vx.voxelsList[i]->pos;   // saveInitialPosition
            }
        }
        currentTime += dt;
        voxcad.afterStep(
            vx.collisions, dtFixed > 0 ? dtFixed : DtFrac * OptimalDt, [&] { for (Voxel* v : vx.voxelsList) v->haltMotion(); },
            [&](bool vel, bool kin, bool strainE, double& maxVel2, double& totKE, double& totSE) {
                for (Voxel* v : vx.voxelsList) {
                    if (vel) { const double v2 = v->velocity().Length2(); if (v2 > maxVel2) maxVel2 = v2; }
                    if (kin) totKE += v->kineticEnergy();
                }
                if (strainE) for (Link* l : vx.linksList) totSE += l->strainEnergy();
            });
        return true;
    }
};

static thread_local std::string g_err;

static bool fail(const std::string& s) { g_err = s; return false; }

/* CVX_Sim::CopyMat (VX_Sim.cpp:368-420) followed by the damping setters of Import (VX_Sim.cpp:81-83) */
static bool copyMat(const vx3h_material_desc& d, MaterialVoxel* p) {
    p->matid = d.matid;
    p->isMeasured = d.is_measured;
    p->thermalOnAfter = d.thermal_on_after;
    p->fixed = d.fixed != 0;
    p->sticky = d.sticky != 0;
    switch (d.mat_model) {
    case VX3H_MDL_LINEAR: p->setModelLinear((float)d.elastic_mod); break;
    case VX3H_MDL_LINEAR_FAIL: p->setModelLinear((float)d.elastic_mod, (float)d.fail_stress); break;
    case VX3H_MDL_BILINEAR: p->setModelBilinear((float)d.elastic_mod, (float)d.plastic_mod, (float)d.yield_stress, (float)d.fail_stress); break;
    case VX3H_MDL_DATA: {
        if (d.n_data <= 0 || d.n_data > VX3H_MAX_CURVE) return fail("bad MDL_DATA point count");
        std::vector<float> st, ss;
        for (int i = 0; i < d.n_data; i++) { ss.push_back((float)d.stress_data[i]); st.push_back((float)d.strain_data[i]); }
        p->setModel(d.n_data, st.data(), ss.data());
        break;
    }
    default: return fail("unknown material model");
    }
    p->setPoissonsRatio((float)d.poissons_ratio);
    p->setDensity((float)d.density);
    p->alphaCTE = (float)d.cte;
    p->setStaticFriction((float)d.u_static);
    p->setKineticFriction((float)d.u_dynamic);
    p->setInternalDamping((float)d.zeta_internal);
    p->setGlobalDamping((float)d.zeta_global);
    p->setCollisionDamping((float)d.zeta_collision);
    return true;
}

static Sim* build(const vx3h_sim_desc* d) {
    if (!d || d->n_voxels < 0 || d->n_materials <= 0) { fail("empty description"); return nullptr; }
    Sim* s = new Sim();
    Engine& vx = s->vx;
    vx.setGravity(d->gravity_mult);
    vx.enableFloor(d->floor_enabled != 0);
    for (int i = 0; i < d->n_materials; i++) {
        MaterialVoxel* p = vx.addMaterial();
        if (!copyMat(d->materials[i], p)) { delete s; return nullptr; }
    }
    vx.setVoxelSize(d->voxel_size);
    for (int i = 0; i < d->n_voxels; i++) {
        int m = d->material[i];
        if (m < 0 || m >= d->n_materials) { fail("voxel material out of range"); delete s; return nullptr; }
        if (vx.voxelAt(d->ijk[3 * i], d->ijk[3 * i + 1], d->ijk[3 * i + 2])) { fail("duplicate voxel"); delete s; return nullptr; }
        Voxel* v = vx.setVoxel(vx.voxelMats[m], d->ijk[3 * i], d->ijk[3 * i + 1], d->ijk[3 * i + 2]);
        v->phaseOffset = d->phase ? d->phase[i] : 0.0;
    }
    // boundary conditions (VX_Sim.cpp:109-143), already resolved per voxel
    for (int e = 0; e < d->n_externals; e++) {
        const vx3h_external& x = d->externals[e];
        if (x.voxel < 0 || x.voxel >= d->n_voxels) { fail("external voxel out of range"); delete s; return nullptr; }
        External* pe = vx.voxelsList[x.voxel]->external();
        const int bits[6] = {X_TRANSLATE, Y_TRANSLATE, Z_TRANSLATE, X_ROTATE, Y_ROTATE, Z_ROTATE};
        for (int k = 0; k < 6; k++)
            if (x.dof_fixed & bits[k]) pe->setDisplacement(bits[k], k < 3 ? x.translate[k] : x.rotate[k - 3]);
        pe->extForce += Vf((float)x.force[0], (float)x.force[1], (float)x.force[2]);
        pe->extMoment += Vf((float)x.moment[0], (float)x.moment[1], (float)x.moment[2]);
    }
    // UpdateMatTemps (VX_Sim.cpp:467-483) with CVX_Environment::UpdateCurTemp at t=0 (VX_Environment.cpp:327-346)
    if (d->temp_enabled) {
        float cur;
        if (d->vary_temp_enabled) {
            if (d->temp_period == 0) cur = 0.0f;
            else cur = (float)(d->temp_base + d->temp_amplitude * sin(2 * 3.1415926 / d->temp_period * 0.0));
        } else cur = (float)(d->temp_base + d->temp_amplitude);
        float t0 = (float)(cur - d->temp_base);
        for (Voxel* v : vx.voxelsList) v->setTemperature(t0);
    }
    // EnableVolumeEffects (VX_Sim.cpp:445-458)
    for (int i = 0; i < d->n_materials; i++)
        vx.voxelMats[i]->setPoissonsRatio(d->volume_effects ? (float)d->materials[i].poissons_ratio : 0.0f);
    if (d->init_velocity)   // starting momentum mass * velocity (no VXA key; synthetic workloads)
        for (size_t i = 0; i < vx.voxelsList.size(); i++) {
            Voxel* v = vx.voxelsList[i];
            v->linMom = Vd(d->init_velocity[3 * i], d->init_velocity[3 * i + 1], d->init_velocity[3 * i + 2]) * (double)v->mat->_mass;
        }
    if (d->collisions_enabled) {
        vx.boundingRadius = (float)d->bounding_radius;
        vx.watchDistance = (float)d->watch_distance;
        vx.enableCollisions(true);
    }
    s->DtFrac = d->dt_frac;
    s->dtFixed = d->dt_fixed;
    s->vary = d->vary_temp_enabled != 0;
    s->enableExpansion = d->enable_expansion != 0;
    s->TempAmplitude = d->temp_amplitude;
    s->TempPeriod = d->temp_period;
    s->maxSteps = d->max_steps;
    // VX3 cilia and force field
    for (int k = 0; k < 3; k++)
        if (d->n_force_field_tokens[k] > 0) s->ffTok[k].assign(d->force_field_tokens[k], d->force_field_tokens[k] + d->n_force_field_tokens[k]);
    if (d->enable_cilia)
        for (int i = 0; i < d->n_materials; i++) s->ciliaOn |= d->materials[i].cilia != 0;
    s->signals = d->enable_signals != 0;
    bool anyTarget = false;
    if (d->collisions_enabled)
        for (size_t i = 0; i < vx.voxelsList.size(); i++)
            if (d->materials[d->material[i]].is_target) { vx.voxelsList[i]->target = true; anyTarget = true; }
    // VX3 order (VX3_VoxelyzeKernel.cu:281-303): links, then updateAttach (target contacts), then
    // the cilia force, then the voxels (with the force field)
    vx.afterContacts = [s, anyTarget] {
        if (anyTarget) s->targetContacts();
        s->updateExtras();
    };
    s->extras = s->ciliaOn || !s->ffTok[0].empty() || !s->ffTok[1].empty() || !s->ffTok[2].empty() || s->signals || anyTarget;
    vx.detach = d->enable_detach != 0;
    vx.attach = d->enable_attach != 0;
    vx.safetyGuard = d->safety_guard;
    if (vx.attach)
        for (int k = 0; k < 5; k++)
            if (d->n_attach_tokens[k] > 0) s->attachTok[k].assign(d->attach_tokens[k], d->attach_tokens[k] + d->n_attach_tokens[k]);
    s->closeR = d->max_dist_close_pairs * d->voxel_size;
    for (size_t i = 0; i < vx.voxelsList.size(); i++)
        if (d->materials[d->material[i]].is_target) s->closeTargets.push_back(vx.voxelsList[i]);
    s->secondary = d->secondary_experiment != 0;
    s->reinitAfter = d->reinit_after;
    for (int i = 0; i < d->n_materials; i++) s->matRemoveAfter.push_back(d->materials[i].remove_after);
    s->matRemoved.assign(d->n_materials, false);
    for (Voxel* v : vx.voxelsList) s->initPos.push_back(v->originalPosition());
    for (int i = 0; i < d->n_materials; i++) { s->matCilia.push_back(d->materials[i].cilia); s->matCiliaOn.push_back(d->materials[i].cilia_on_after); }
    for (size_t i = 0; i < vx.voxelsList.size(); i++) {
        Voxel* v = vx.voxelsList[i];
        v->extras = s->extras;
        if (d->base_cilia_force) v->baseCilia = Vd(d->base_cilia_force[3 * i], d->base_cilia_force[3 * i + 1], d->base_cilia_force[3 * i + 2]);
        if (d->shift_cilia_force) v->shiftCilia = Vd(d->shift_cilia_force[3 * i], d->shift_cilia_force[3 * i + 1], d->shift_cilia_force[3 * i + 2]);
        const vx3h_material_desc& md = d->materials[d->material[i]];
        v->isPaceMaker = md.is_pacemaker != 0;
        v->paceMakerPeriod = md.pacemaker_period;
        v->signalValueDecay = md.signal_value_decay;
        v->signalTimeDelay = md.signal_time_delay;
        v->inactivePeriod = md.inactive_period;
    }
    s->voxcad.type = d->stop_condition_type;
    s->voxcad.value = d->stop_condition_value;
    s->voxcad.eq = d->equilibrium_mode != 0;
    if (d->n_stop_tokens > 0) s->stopTok.assign(d->stop_tokens, d->stop_tokens + d->n_stop_tokens);
    if (d->n_fitness_tokens > 0) s->fitTok.assign(d->fitness_tokens, d->fitness_tokens + d->n_fitness_tokens);
    // VX3_SimulationManager.cu:54-55
    s->updateCurrentCenterOfMass();
    s->initialCoM = s->currentCoM;
    return s;
}

}  // namespace vxo

using namespace vxo;
struct vxo_sim { Sim* s; };

extern "C" {

const char* vxo_last_error(void) { return g_err.c_str(); }

vxo_sim* vxo_create(const vx3h_sim_desc* desc) {
    Sim* s = build(desc);
    if (!s) return nullptr;
    vxo_sim* h = new vxo_sim;
    h->s = s;
    return h;
}
void vxo_free(vxo_sim* h) { if (h) { delete h->s; delete h; } }
void vxo_counts(vxo_sim* h, int* nv, int* nl) {
    if (nv) *nv = (int)h->s->vx.voxelsList.size();
    if (nl) *nl = (int)h->s->vx.linksList.size();
}
void vxo_topology(vxo_sim* h, int* vneg, int* vpos, int* axis, int* voxel_links) {
    Engine& e = h->s->vx;
    std::map<Link*, int> li;
    for (size_t i = 0; i < e.linksList.size(); i++) {
        Link* l = e.linksList[i];
        li[l] = (int)i;
        if (vneg) vneg[i] = l->pVNeg->index;
        if (vpos) vpos[i] = l->pVPos->index;
        if (axis) axis[i] = l->axis;
    }
    if (voxel_links)
        for (size_t v = 0; v < e.voxelsList.size(); v++)
            for (int d = 0; d < 6; d++) voxel_links[6 * v + d] = e.voxelsList[v]->links[d] ? li[e.voxelsList[v]->links[d]] : -1;
}
int vxo_do_time_step(vxo_sim* h, float dt) { return h->s->vx.doTimeStep(dt) ? 1 : 0; }
float vxo_recommended_dt(vxo_sim* h) { return h->s->vx.recommendedTimeStep(); }
void vxo_reset_time(vxo_sim* h) { h->s->vx.resetTime(); }
void vxo_set_gravity(vxo_sim* h, float g) { h->s->vx.setGravity(g); }
void vxo_enable_floor(vxo_sim* h, int en) { h->s->vx.enableFloor(en != 0); }
void vxo_enable_collisions(vxo_sim* h, int en) { h->s->vx.enableCollisions(en != 0); }
void vxo_set_ambient_temperature(vxo_sim* h, float t, int all) { h->s->vx.setAmbientTemperature(t, all != 0); }
void vxo_set_voxel_temperature(vxo_sim* h, int v, float t) { h->s->vx.voxelsList[v]->setTemperature(t); }
void vxo_ext_set_force(vxo_sim* h, int v, float fx, float fy, float fz) { h->s->vx.voxelsList[v]->external()->extForce = Vf(fx, fy, fz); }
void vxo_ext_set_moment(vxo_sim* h, int v, float mx, float my, float mz) { h->s->vx.voxelsList[v]->external()->extMoment = Vf(mx, my, mz); }
void vxo_ext_set_fixed(vxo_sim* h, int v, int m) {
    h->s->vx.voxelsList[v]->external()->setFixed(m & 1, m & 2, m & 4, m & 8, m & 16, m & 32);
}
void vxo_ext_set_fixed_all(vxo_sim* h, int v, int fixed) {
    External* e = h->s->vx.voxelsList[v]->external();
    if (fixed) e->setDisplacementAll(Vd(0, 0, 0), Vd(0, 0, 0));
    else e->clearDisplacementAll();
}
void vxo_ext_set_displacement_all(vxo_sim* h, int v, double tx, double ty, double tz, double rx, double ry, double rz) {
    h->s->vx.voxelsList[v]->external()->setDisplacementAll(Vd(tx, ty, tz), Vd(rx, ry, rz));
}
void vxo_external_force(vxo_sim* h, int v, float* out) {
    Voxel* pV = h->s->vx.voxelsList[v];
    External* ext = pV->external();
    Vf r = ext->extForce;
    if (ext->isFixed(X_TRANSLATE) || ext->isFixed(Y_TRANSLATE) || ext->isFixed(Z_TRANSLATE)) {
        Vf t = (Vf)(-pV->force());
        if (ext->isFixed(X_TRANSLATE)) r.x = t.x;
        if (ext->isFixed(Y_TRANSLATE)) r.y = t.y;
        if (ext->isFixed(Z_TRANSLATE)) r.z = t.z;
    }
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void vxo_material_set(vxo_sim* h, int m, int which, float value) {
    MaterialVoxel* p = h->s->vx.voxelMats[m];
    switch (which) {
    case 0: p->setInternalDamping(value); break;
    case 1: p->setGlobalDamping(value); break;
    case 2: p->setCollisionDamping(value); break;
    case 3: p->setStaticFriction(value); break;
    case 4: p->setKineticFriction(value); break;
    case 5: p->alphaCTE = value; break;
    case 6: p->setPoissonsRatio(value); break;
    default: break;
    }
}
void vxo_replace_voxel_material(vxo_sim* h, int v, int m) {
    Voxel* pV = h->s->vx.voxelsList[v];
    h->s->vx.setVoxel(h->s->vx.voxelMats[m], pV->ix, pV->iy, pV->iz);
}
int vxo_vx3_step(vxo_sim* h, int n) {
    Sim* s = h->s;
    int done = 0;
    for (int j = 0; j < n; j++) {
        if (s->state != VX3H_SIM_RUNNING) break;
        if (s->stopConditionMet()) { s->state = VX3H_SIM_DONE; break; }
        if (!s->doTimeStepVX3()) { s->state = VX3H_SIM_DIVERGED; break; }
        done++;
    }
    if (s->state == VX3H_SIM_RUNNING && s->stopConditionMet()) s->state = VX3H_SIM_DONE;
    return done;
}
void vxo_vx3_status(vxo_sim* h, vx3h_sim_status* st) {
    Sim* s = h->s;
    Vd saveCur = s->currentCoM;
    s->updateCurrentCenterOfMass();
    Vd off = s->currentCoM - s->initialCoM;
    memset(st, 0, sizeof(*st));
    st->state = s->state;
    st->steps = (int)s->CurStepCount;
    st->num_voxel = (int)s->vx.voxelsList.size();
    st->current_time = s->currentTime;
    st->dt = s->lastDt;
    st->recommended_dt = s->OptimalDt;
    st->collision_count = s->hits;
    st->initial_com[0] = s->initialCoM.x; st->initial_com[1] = s->initialCoM.y; st->initial_com[2] = s->initialCoM.z;
    st->current_com[0] = s->currentCoM.x; st->current_com[1] = s->currentCoM.y; st->current_com[2] = s->currentCoM.z;
    st->num_close_pairs = s->numClosePairs;
    st->target_closeness = s->targetCloseness;
    st->fitness = mathtree_eval(s->fitTok.data(), (int)s->fitTok.size(), off.x, off.y, off.z, s->hits, s->currentTime, s->recentAngle,
                                s->targetCloseness, s->numClosePairs,
                                (int)s->vx.voxelsList.size());
    st->recent_angle = s->recentAngle;
    int nm = 0;
    double tot = 0;
    for (size_t i = 0; i < s->vx.voxelsList.size(); i++) {
        Voxel* v = s->vx.voxelsList[i];
        if (!v->mat->isMeasured) continue;
        nm++;
        tot += Vd(v->pos.x, v->pos.y, v->pos.z).Dist(s->initPos[i]);   // d_initialPosition (VX3_SimulationManager.cu:439-452)
    }
    st->num_measured_voxel = nm;
    st->total_distance_of_all_voxels = tot;
    s->currentCoM = saveCur;
}
void vxo_voxel_state(vxo_sim* h, double* pos, double* orient, double* lm, double* am, int* flags, float* temp) {
    Engine& e = h->s->vx;
    for (size_t i = 0; i < e.voxelsList.size(); i++) {
        Voxel* v = e.voxelsList[i];
        if (pos) { pos[3 * i] = v->pos.x; pos[3 * i + 1] = v->pos.y; pos[3 * i + 2] = v->pos.z; }
        if (orient) { orient[4 * i] = v->orient.w; orient[4 * i + 1] = v->orient.x; orient[4 * i + 2] = v->orient.y; orient[4 * i + 3] = v->orient.z; }
        if (lm) { lm[3 * i] = v->linMom.x; lm[3 * i + 1] = v->linMom.y; lm[3 * i + 2] = v->linMom.z; }
        if (am) { am[3 * i] = v->angMom.x; am[3 * i + 1] = v->angMom.y; am[3 * i + 2] = v->angMom.z; }
        if (flags) flags[i] = v->boolStates | (v->removed ? 16 : 0);
        if (temp) temp[i] = v->temp;
    }
}
void vxo_link_state(vxo_sim* h, float* strain, float* maxStrain, float* off, int* sa, int* bs, double* rl, double* f12) {
    Engine& e = h->s->vx;
    for (size_t i = 0; i < e.linksList.size(); i++) {
        Link* l = e.linksList[i];
        if (strain) strain[i] = l->strain;
        if (maxStrain) maxStrain[i] = l->maxStrain;
        if (off) off[i] = l->strainOffset;
        if (sa) sa[i] = l->smallAngle ? 1 : 0;
        if (bs) bs[i] = l->boolStates;
        if (rl) rl[i] = l->currentRestLength;
        if (f12) {
            const Vd* vs[4] = {&l->forceNeg, &l->forcePos, &l->momentNeg, &l->momentPos};
            for (int k = 0; k < 4; k++) { f12[12 * i + 3 * k] = vs[k]->x; f12[12 * i + 3 * k + 1] = vs[k]->y; f12[12 * i + 3 * k + 2] = vs[k]->z; }
        }
    }
}
void vxo_local_signal(vxo_sim* h, double* out) {
    for (size_t i = 0; i < h->s->vx.voxelsList.size(); i++) out[i] = h->s->vx.voxelsList[i]->localSignal;
}
void vxo_voxcad_history(vxo_sim* h, double* k, double* t, double* m) {
    const VoxcadStop& c = h->s->voxcad;
    for (int i = 0; i < VoxcadStop::HIST_LEN; i++) { k[i] = c.KinEHistory[i]; t[i] = c.TotEHistory[i]; m[i] = c.MaxMoveHistory[i]; }
}
void vxo_corner_offset(vxo_sim* h, int v, int corner, float* out) {
    Vf r = h->s->vx.voxelsList[v]->cornerOffset(corner);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void vxo_set_collision_interval(vxo_sim* h, int n) { h->s->vx.colInterval = n > 0 ? n : 0; }
int vxo_collision_pairs(vxo_sim* h, int* pairs, int cap) {
    Engine& e = h->s->vx;
    int n = (int)e.collisionsList.size();
    for (int i = 0; i < n && i < cap; i++) { pairs[2 * i] = e.collisionsList[i]->pV1->index; pairs[2 * i + 1] = e.collisionsList[i]->pV2->index; }
    return n;
}
int vxo_voxel_material_consts(vxo_sim* h, float* out, int cap) {
    Engine& e = h->s->vx;
    int n = (int)e.voxelMats.size();
    for (int i = 0; i < n && i < cap; i++) {
        MaterialVoxel* m = e.voxelMats[i];
        float* o = out + 14 * i;
        o[0] = m->_mass; o[1] = m->_massInverse; o[2] = m->_sqrtMass; o[3] = m->_momentInertia; o[4] = m->_momentInertiaInverse;
        o[5] = m->_2xSqMxExS; o[6] = m->_2xSqIxExSxSxS; o[7] = m->E; o[8] = m->nu; o[9] = m->_eHat; o[10] = m->penetrationStiffness();
        o[11] = m->gravityForce(); o[12] = m->muStatic; o[13] = m->muKinetic;
    }
    return n;
}
int vxo_link_material_consts(vxo_sim* h, int* link_mat_index, float* out, int cap) {
    Engine& e = h->s->vx;
    std::vector<MaterialLink*> mats(e.linkMats.begin(), e.linkMats.end());
    if (link_mat_index)
        for (size_t i = 0; i < e.linksList.size(); i++)
            link_mat_index[i] = (int)(std::find(mats.begin(), mats.end(), e.linksList[i]->mat) - mats.begin());
    int n = (int)mats.size();
    for (int i = 0; i < n && i < cap; i++) {
        MaterialLink* m = mats[i];
        float* o = out + 20 * i;
        o[0] = m->_a1; o[1] = m->_a2; o[2] = m->_b1; o[3] = m->_b2; o[4] = m->_b3; o[5] = m->_sqA1; o[6] = m->_sqA2xIp; o[7] = m->_sqB1;
        o[8] = m->_sqB2xFMp; o[9] = m->_sqB3xIp; o[10] = m->E; o[11] = m->nu; o[12] = m->_eHat; o[13] = m->linear ? 1.f : 0.f;
        o[14] = m->epsilonFail; o[15] = m->epsilonYield; o[16] = m->alphaCTE; o[17] = (float)m->strainData.size();
        o[18] = m->strainData.size() > 1 ? m->strainData[1] : 0; o[19] = m->stressData.size() > 1 ? m->stressData[1] : 0;
    }
    return n;
}
double vxo_mathtree_eval(const vx3h_token* toks, int n, double x, double y, double z, double hit, double t, double angle, double closeness,
                         int ncp, int nv) {
    return mathtree_eval(toks, n, x, y, z, hit, t, angle, closeness, ncp, nv);
}

}  // extern "C"
