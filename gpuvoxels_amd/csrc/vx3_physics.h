// vx3_physics.h -- device physics of one link and one voxel, templated on the
// state precision T.  With T = double every operation matches the CPU Voxelyze
// engine's precision and order (positions/orientations in double, strains and
// material constants in float) so the double instantiation reproduces the CPU
// oracle to rounding of the transcendental functions; with T = float it is the
// FP32 product path.
//
// Reference semantics (paths relative to the gpuVoxels tree):
//   Quat3D ops               Voxelyze3/src/old/Quat3D.h:87-208
//   CVX_Link::orientLink     Voxelyze3/src/old/VX_Link.cpp:72-113
//   CVX_Link::updateForces   Voxelyze3/src/old/VX_Link.cpp:143-209
//   CVX_Link::updateStrain   Voxelyze3/src/old/VX_Link.cpp:212-241
//   CVX_Material::stress     Voxelyze3/src/old/VX_Material.cpp:182-212
//   CVX_Voxel::timeStep      Voxelyze3/src/old/VX_Voxel.cpp:163-230
//   CVX_Voxel::floorForce    Voxelyze3/src/old/VX_Voxel.cpp:272-297
//   CVX_Voxel::strain        Voxelyze3/src/old/VX_Voxel.cpp:299-333
#ifndef VX3_PHYSICS_H
#define VX3_PHYSICS_H

#include <hip/hip_runtime.h>
#include "vx3_types.h"

namespace vx3 {

__device__ __forceinline__ float msqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double msqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float macos(float x) { return acosf(x); }
__device__ __forceinline__ double macos(double x) { return acos(x); }
__device__ __forceinline__ float msin(float x) { return sinf(x); }
__device__ __forceinline__ double msin(double x) { return sin(x); }
__device__ __forceinline__ float mcos(float x) { return cosf(x); }
__device__ __forceinline__ double mcos(double x) { return cos(x); }
__device__ __forceinline__ float mabs(float x) { return fabsf(x); }
__device__ __forceinline__ double mabs(double x) { return fabs(x); }

template <typename T> struct V3 {
    T x, y, z;
    __device__ V3() : x(0), y(0), z(0) {}
    __device__ V3(T a, T b, T c) : x(a), y(b), z(c) {}
    __device__ V3 operator+(const V3& v) const { return V3(x + v.x, y + v.y, z + v.z); }
    __device__ V3 operator-(const V3& v) const { return V3(x - v.x, y - v.y, z - v.z); }
    __device__ V3 operator-() const { return V3(-x, -y, -z); }
    __device__ V3 mul(T f) const { return V3(f * x, f * y, f * z); }
    __device__ V3 div(T f) const { T inv = (T)1.0 / f; return V3(inv * x, inv * y, inv * z); }
    __device__ void add(const V3& s) { x += s.x; y += s.y; z += s.z; }
    __device__ void sub(const V3& s) { x -= s.x; y -= s.y; z -= s.z; }
    __device__ T len2() const { return x * x + y * y + z * z; }
    __device__ T len() const { return msqrt(x * x + y * y + z * z); }
    __device__ T dot(const V3& v) const { return x * v.x + y * v.y + z * v.z; }
    __device__ V3 normalized() const { T l = msqrt(x * x + y * y + z * z); return l > 0 ? div(l) : *this; }
};

template <typename T> struct Q4 {
    T w, x, y, z;
    __device__ Q4() : w(1), x(0), y(0), z(0) {}
    __device__ Q4(T a, T b, T c, T d) : w(a), x(b), y(c), z(d) {}
    __device__ Q4 operator*(const Q4& f) const {
        return Q4(w * f.w - x * f.x - y * f.y - z * f.z, w * f.x + x * f.w + y * f.z - z * f.y,
                  w * f.y - x * f.z + y * f.w + z * f.x, w * f.z + x * f.y - y * f.x + z * f.w);
    }
    __device__ Q4 conj() const { return Q4(w, -x, -y, -z); }
    // Quat3D::ToRotationVector (Quat3D.h:121-126).  halfAngle >= 0 is acos(w) when
    // the caller knows it (the quaternion came from fromAngleToPosX).
    __device__ V3<T> toRotVec(T halfAngle = (T)-1) const {
        if (w >= (T)1.0 || w <= (T)-1.0) return V3<T>(0, 0, 0);
        T sl = (T)1.0 - w * w;
        if (sl < (T)2.4e-3) { T f = msqrt(((T)2 - (T)2 * w) / sl); return V3<T>(f * ((T)2.0 * x), f * ((T)2.0 * y), f * ((T)2.0 * z)); }
        T a = halfAngle >= 0 ? halfAngle : macos(w), inv = (T)1.0 / msqrt(sl);
        return V3<T>(inv * (a * ((T)2.0 * x)), inv * (a * ((T)2.0 * y)), inv * (a * ((T)2.0 * z)));
    }
    // Quat3D::FromRotationVector (Quat3D.h:128-143)
    __device__ static Q4 fromRotVec(const V3<T>& v) {
        V3<T> th = v.div((T)2);
        T s, m2 = th.len2();
        Q4 q;
        if (m2 * m2 < (T)5.328e-15) { q.w = (T)1.0 - (T)0.5 * m2; s = (T)1.0 - m2 / (T)6.0; }
        else { T m = msqrt(m2); q.w = mcos(m); s = msin(m) / m; }
        q.x = th.x * s; q.y = th.y * s; q.z = th.z * s;
        return q;
    }
    // Quat3D::FromAngleToPosX (Quat3D.h:145-170).  Returns theta/2 when the general
    // branch ran (for toRotVec below), -1 otherwise.  The half-angle cosine and sine
    // come from cos(theta) = n.x by the half-angle identities, cos(theta/2) =
    // sqrt((1 + n.x)/2) and sin(theta/2) = sqrt((1 - n.x)/2) (theta in [0, pi]),
    // instead of cos/sin of acos(n.x): equal to within rounding, and two square
    // roots in place of two trigonometric calls.
    __device__ T fromAngleToPosX(const V3<T>& r) {
        if (r.x == 0 && r.y == 0 && r.z == 0) return (T)-1;
        T yx = r.y / r.x, zx = r.z / r.x;
        const T SA = (T)1.732e-2;
        if (yx < SA && yx > -SA && zx < SA && zx > -SA) {
            x = 0; y = (T)0.5 * zx; z = (T)-0.5 * yx;
            w = (T)1 + (T)0.5 * (-y * y - z * z);
            return (T)-1;
        }
        V3<T> n = r;
        { T l = msqrt(n.x * n.x + n.y * n.y + n.z * n.z); if (l > 0) { T li = (T)1.0 / l; n.x *= li; n.y *= li; n.z *= li; } }
        T theta = macos(n.x);
        if (theta > (T)(3.14159265358979 - 1e-7)) { w = 0; x = 0; y = 1; z = 0; return (T)-1; }
        const T ami = (T)1.0 / msqrt(n.z * n.z + n.y * n.y);
        const T s = msqrt(((T)1 - n.x) * (T)0.5);
        w = msqrt(((T)1 + n.x) * (T)0.5); x = 0; y = n.z * ami * s; z = -n.y * ami * s;
        return (T)0.5 * theta;
    }
    __device__ V3<T> rot(const V3<T>& f) const {   // RotateVec3D
        T fx = f.x, fy = f.y, fz = f.z;
        T tw = fx * x + fy * y + fz * z;
        T tx = fx * w - fy * z + fz * y;
        T ty = fx * z + fy * w - fz * x;
        T tz = -fx * y + fy * x + fz * w;
        return V3<T>(w * tx + x * tw + y * tz - z * ty, w * ty - x * tz + y * tw + z * tx, w * tz + x * ty - y * tx + z * tw);
    }
    __device__ V3<T> rotInv(const V3<T>& f) const {   // RotateVec3DInv
        T fx = f.x, fy = f.y, fz = f.z;
        T tw = x * fx + y * fy + z * fz;
        T tx = w * fx - y * fz + z * fy;
        T ty = w * fy + x * fz - z * fx;
        T tz = w * fz - x * fy + y * fx;
        return V3<T>(tw * x + tx * w + ty * z - tz * y, tw * y - tx * z + ty * w + tz * x, tw * z + tx * y - ty * x + tz * w);
    }
};

// CVX_Link::toAxisX / toAxisOriginal (VX_Link.h:115-120)
template <typename T> __device__ __forceinline__ V3<T> toAxisX(int axis, const V3<T>& v) {
    return axis == 1 ? V3<T>(v.y, -v.x, v.z) : (axis == 2 ? V3<T>(v.z, v.y, -v.x) : v);
}
template <typename T> __device__ __forceinline__ Q4<T> toAxisX(int axis, const Q4<T>& q) {
    return axis == 1 ? Q4<T>(q.w, q.y, -q.x, q.z) : (axis == 2 ? Q4<T>(q.w, q.z, q.y, -q.x) : q);
}
template <typename T> __device__ __forceinline__ V3<T> toAxisOriginal(int axis, const V3<T>& v) {
    return axis == 1 ? V3<T>(-v.y, v.x, v.z) : (axis == 2 ? V3<T>(-v.z, v.y, v.x) : v);
}

// CVX_Material::stress for link materials (VX_Material.cpp:182-212)
__device__ __forceinline__ float matStress(const LinkMatDev& m, float strain, float tss, bool forceLinear) {
    if (m.epsilonFail != -1.0f && strain > m.epsilonFail) return 0.0f;
    if (strain <= m.strainData[1] || m.linear || forceLinear) {
        if (m.nu == 0.0f) return m.E * strain;
        return m.eHat * ((1 - m.nu) * strain + m.nu * tss);
    }
    const int n = m.nData;
    for (int i = 2; i < n; i++) {
        if (strain <= m.strainData[i] || i == n - 1) {
            float perc = (strain - m.strainData[i - 1]) / (m.strainData[i] - m.strainData[i - 1]);
            float basic = m.stressData[i - 1] + perc * (m.stressData[i] - m.stressData[i - 1]);
            if (m.nu == 0.0f) return basic;
            float mod = (m.stressData[i] - m.stressData[i - 1]) / (m.strainData[i] - m.strainData[i - 1]);
            float modHat = mod / ((1 - 2 * m.nu) * (1 + m.nu));
            float effStrain = basic / mod;
            float effTss = tss * (effStrain / strain);
            return modHat * ((1 - m.nu) * effStrain + m.nu * effTss);
        }
    }
    return 0.0f;
}

// Mutable per-link state.
template <typename T> struct LinkState {
    V3<T> pos2, a1v, a2v;
    float strain, maxStrain, strainOffset;
    int flags;   // LF_*
};

// One beam element: CVX_Link::updateForces.  Returns the four local end loads;
// st is updated in place.  dmNeg/dmPos are the voxels' damping multipliers
// (2*sqrt(m)*zeta/previousDt, VX_Voxel.h:131).
template <typename T>
__device__ __forceinline__ void linkForces(const LinkMatDev& mat, int axis, const V3<T>& pNeg, const Q4<T>& oNeg,
                                           const V3<T>& pPos, const Q4<T>& oPos, T restLength, float area, float tss,
                                           bool volumeLink, float dmNeg, float dmPos, LinkState<T>& st, V3<T>& fNeg,
                                           V3<T>& mNeg, V3<T>& fPos, V3<T>& mPos) {
    const V3<T> oldPos2 = st.pos2, oldA1 = st.a1v, oldA2 = st.a2v;
    // orientLink
    V3<T> pos2 = toAxisX(axis, pPos - pNeg);
    Q4<T> angle1 = toAxisX(axis, oNeg);
    Q4<T> angle2 = toAxisX(axis, oPos);
    Q4<T> totalRot = angle1.conj();
    pos2 = totalRot.rot(pos2);
    angle2 = totalRot * angle2;
    angle1 = Q4<T>();
    float smallTurn = (float)((mabs(pos2.z) + mabs(pos2.y)) / pos2.x);
    float extendPerc = (float)(mabs((T)1 - pos2.x / restLength));
    bool small = (st.flags & LF_SMALL_ANGLE) != 0;
    T halfTheta = (T)-1;
    if (!small && smallTurn < 0.05f && extendPerc < 0.50f) {
        small = true;
        st.flags &= ~LF_LOCAL_VELOCITY_VALID;
    } else if (small && (smallTurn > 1.2f * 0.05f || extendPerc > 1.2f * 0.50f)) {
        small = false;
        st.flags &= ~LF_LOCAL_VELOCITY_VALID;
    }
    if (small) {
        st.flags |= LF_SMALL_ANGLE;
        pos2.x -= restLength;
    } else {
        st.flags &= ~LF_SMALL_ANGLE;
        halfTheta = angle1.fromAngleToPosX(pos2);
        angle2 = angle1 * angle2;
        pos2 = V3<T>(pos2.len() - restLength, 0, 0);
    }
    const V3<T> a1v = angle1.toRotVec(halfTheta);
    const V3<T> a2v = angle2.toRotVec();
    st.pos2 = pos2; st.a1v = a1v; st.a2v = a2v;

    const V3<T> dPos2 = (pos2 - oldPos2).mul((T)0.5);
    const V3<T> dA1 = (a1v - oldA1).mul((T)0.5);
    const V3<T> dA2 = (a2v - oldA2).mul((T)0.5);

    // updateStrain (VX_Link.cpp:212-241)
    const float axialStrain = (float)(pos2.x / restLength);
    st.strain = axialStrain;
    float stress;
    if (mat.linear) {
        if (axialStrain > st.maxStrain) st.maxStrain = axialStrain;
        stress = matStress(mat, axialStrain, tss, false);
    } else if (axialStrain > st.maxStrain) {
        st.maxStrain = axialStrain;
        stress = matStress(mat, axialStrain, tss, false);
        if (mat.nu != 0.0f) st.strainOffset = st.maxStrain - matStress(mat, axialStrain, 0.0f, false) / (mat.eHat * (1 - mat.nu));
        else st.strainOffset = st.maxStrain - stress / mat.E;
    } else {
        float rel = axialStrain - st.strainOffset;
        if (mat.nu != 0.0f) stress = matStress(mat, rel, tss, true);
        else stress = mat.E * rel;
    }
    (void)volumeLink;
    if (mat.epsilonFail != -1.0f && st.maxStrain > mat.epsilonFail) {
        fNeg = fPos = mNeg = mPos = V3<T>(0, 0, 0);
        return;
    }
    const float b1 = mat.b1, b2 = mat.b2, b3 = mat.b3, a2 = mat.a2;
    fNeg = V3<T>((T)(stress * area), (T)b1 * pos2.y - (T)b2 * (a1v.z + a2v.z), (T)b1 * pos2.z + (T)b2 * (a1v.y + a2v.y));
    fPos = -fNeg;
    mNeg = V3<T>((T)a2 * (a2v.x - a1v.x), (T)(-b2) * pos2.z - (T)b3 * ((T)2 * a1v.y + a2v.y),
                 (T)b2 * pos2.y - (T)b3 * ((T)2 * a1v.z + a2v.z));
    mPos = V3<T>((T)a2 * (a1v.x - a2v.x), (T)(-b2) * pos2.z - (T)b3 * (a1v.y + (T)2 * a2v.y),
                 (T)b2 * pos2.y - (T)b3 * (a1v.z + (T)2 * a2v.z));
    if (st.flags & LF_LOCAL_VELOCITY_VALID) {
        const T sqA1 = mat.sqA1, sqA2xIp = mat.sqA2xIp, sqB1 = mat.sqB1, sqB2xFMp = mat.sqB2xFMp, sqB3xIp = mat.sqB3xIp;
        V3<T> posCalc(sqA1 * dPos2.x, sqB1 * dPos2.y - sqB2xFMp * (dA1.z + dA2.z), sqB1 * dPos2.z + sqB2xFMp * (dA1.y + dA2.y));
        fNeg.add(posCalc.mul((T)dmNeg));
        fPos.sub(posCalc.mul((T)dmPos));
        mNeg.sub(V3<T>(-sqA2xIp * (dA2.x - dA1.x), sqB2xFMp * dPos2.z + sqB3xIp * ((T)2 * dA1.y + dA2.y),
                       -sqB2xFMp * dPos2.y + sqB3xIp * ((T)2 * dA1.z + dA2.z)).mul((T)0.5 * (T)dmNeg));
        mPos.sub(V3<T>(sqA2xIp * (dA2.x - dA1.x), sqB2xFMp * dPos2.z + sqB3xIp * (dA1.y + (T)2 * dA2.y),
                       -sqB2xFMp * dPos2.y + sqB3xIp * (dA1.z + (T)2 * dA2.z)).mul((T)0.5 * (T)dmPos));
    } else {
        st.flags |= LF_LOCAL_VELOCITY_VALID;
    }
    if (!small) {
        fNeg = angle1.rotInv(fNeg);
        mNeg = angle1.rotInv(mNeg);
    }
    fPos = angle2.rotInv(fPos);
    mPos = angle2.rotInv(mPos);
    fNeg = toAxisOriginal(axis, fNeg);
    fPos = toAxisOriginal(axis, fPos);
    mNeg = toAxisOriginal(axis, mNeg);
    mPos = toAxisOriginal(axis, mPos);
}

// No collision contacts (the resident path, or a simulation without collisions).
struct NoContacts {
    template <typename T> __device__ __forceinline__ void subtract(V3<T>&) const {}
};

// Voxel translation + rotation update: CVX_Voxel::timeStep after force() and
// moment() have summed the link loads (fLocal/mLocal, voxel frame).  contacts
// subtracts the contact force of each watched pair from the global force, one by
// one in the voxel's colWatch order (VX_Voxel.cpp:245-249).
template <typename T, typename Contacts>
__device__ __forceinline__ void voxelStep(const VoxMatDev& m, const ExtDev* ext, float dt, T nomSize, int ix, int iy, int iz,
                                          float tfac, const V3<T>& fLocal, const V3<T>& mLocal, const Contacts& contacts,
                                          V3<T>& pos, V3<T>& linMom, Q4<T>& orient, V3<T>& angMom, int& flags) {
    if (ext && (ext->dofFixed & 0x3F) == 0x3F) {
        pos = V3<T>((T)(ix * (double)nomSize + ext->translate[0]), (T)(iy * (double)nomSize + ext->translate[1]),
                    (T)(iz * (double)nomSize + ext->translate[2]));
        orient = ext->hasRotQ ? Q4<T>((T)ext->rotQ[0], (T)ext->rotQ[1], (T)ext->rotQ[2], (T)ext->rotQ[3]) : Q4<T>();
        linMom = angMom = V3<T>(0, 0, 0);
        return;
    }
    // force() (VX_Voxel.cpp:232-251)
    V3<T> curForce = orient.rot(fLocal);
    if (ext) curForce.add(V3<T>((T)ext->force[0], (T)ext->force[1], (T)ext->force[2]));
    curForce.sub(linMom.mul((T)m.massInverse).mul((T)(m.zetaGlobal * m.c2xSqMxExS)));
    curForce.z += (T)m.gravityForce;
    contacts.subtract(curForce);
    V3<T> fricForce = curForce;
    const bool floorOn = (flags & VF_FLOOR) != 0;
    // floorPenetration (VX_Voxel.h:124): baseSizeAverage()/2 - nomSize/2 - pos.z
    const T bs = nomSize * (T)tfac;
    const float pen = (float)((bs + bs + bs) / (T)3.0f / (T)2 - nomSize / (T)2 - pos.z);
    if (floorOn) {
        // floorForce (VX_Voxel.cpp:272-297)
        if (pen >= 0) {
            V3<T> vel = linMom.mul((T)m.massInverse);
            V3<T> hv(vel.x, vel.y, 0);
            float normalForce = m.penetrationStiffness * pen;
            curForce.z += (T)normalForce - (T)(m.zetaCollision * m.c2xSqMxExS) * vel.z;
            if (flags & VF_STATIC) {
                float sfs = (float)(curForce.x * curForce.x + curForce.y * curForce.y);
                float ffs = (m.muStatic * normalForce) * (m.muStatic * normalForce);
                if (sfs > ffs) flags &= ~VF_STATIC;
            } else {
                curForce.sub(hv.normalized().mul((T)(m.muKinetic * normalForce)));
            }
        } else flags &= ~VF_STATIC;
    }
    fricForce = curForce - fricForce;
    linMom.add(curForce.mul((T)dt));
    V3<T> translate = linMom.mul((T)(dt * m.massInverse));
    if (floorOn && pen >= 0) {
        T work = fricForce.x * translate.x + fricForce.y * translate.y;
        T hKe = (T)0.5 * (T)m.massInverse * (linMom.x * linMom.x + linMom.y * linMom.y);
        if (hKe + work <= 0) flags |= VF_STATIC;
        if (flags & VF_STATIC) {
            linMom.x = linMom.y = 0;
            translate.x = translate.y = 0;
        }
    } else flags &= ~VF_STATIC;
    pos.add(translate);
    // moment() (VX_Voxel.cpp:253-269)
    V3<T> curMoment = orient.rot(mLocal);
    if (ext) curMoment.add(V3<T>((T)ext->moment[0], (T)ext->moment[1], (T)ext->moment[2]));
    curMoment.sub(angMom.mul((T)m.momentInertiaInverse).mul((T)(m.zetaGlobal * m.c2xSqIxExSxSxS)));
    angMom.add(curMoment.mul((T)dt));
    orient = Q4<T>::fromRotVec(angMom.mul((T)(dt * m.momentInertiaInverse))) * orient;
    if (ext) {
        const int d = ext->dofFixed;
        if (d & 1) { pos.x = (T)(ix * (double)nomSize + ext->translate[0]); linMom.x = 0; }
        if (d & 2) { pos.y = (T)(iy * (double)nomSize + ext->translate[1]); linMom.y = 0; }
        if (d & 4) { pos.z = (T)(iz * (double)nomSize + ext->translate[2]); linMom.z = 0; }
        if (d & (8 | 16 | 32)) {
            if ((d & (8 | 16 | 32)) == (8 | 16 | 32)) {
                orient = ext->hasRotQ ? Q4<T>((T)ext->rotQ[0], (T)ext->rotQ[1], (T)ext->rotQ[2], (T)ext->rotQ[3]) : Q4<T>();
                angMom = V3<T>(0, 0, 0);
            } else {
                V3<T> rv = orient.toRotVec();
                if (d & 8) { rv.x = 0; angMom.x = 0; }
                if (d & 16) { rv.y = 0; angMom.y = 0; }
                if (d & 32) { rv.z = 0; angMom.z = 0; }
                orient = Q4<T>::fromRotVec(rv);
            }
        }
    }
}

}  // namespace vx3
#endif
