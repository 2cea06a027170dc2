/*
 * ref_driver.cpp -- exposes the reference's own CPU engine through the oracle
 * C API (vx_oracle.h).  TEST INFRASTRUCTURE ONLY.
 *
 * Built by oracle/Makefile against the reference sources where they lie under
 * /root/reference (nothing is copied); the result goes to oracle/_ref/.  The
 * physics is the reference's CVoxelyze (Voxelyze3/src/old/); this file only
 *  - builds a CVoxelyze from a vx3h_sim_desc with the same call sequence as
 *    CVX_Sim::Import (Voxelyze3/src/VXA/VX_Sim.cpp:59-153), or loads a VXA text
 *    through the reference's own CVX_Sim::ReadVXA + Import (ref_load_vxa), and
 *  - drives it with the VX3 step semantics (per-voxel phase temperature,
 *    cached dt, measured-voxel center of mass) restated from
 *    Voxelyze3/src/VX3/VX3_VoxelyzeKernel.cu:184-359,477-534,625-650.
 */
#include "vx_oracle.h"
#include "vx_voxcad_stop.h"

#include "Voxelyze.h"
#include "VX_Collision.h"
#include "VX_External.h"
#include "VX_Link.h"
#include "VX_MaterialLink.h"
#include "VX_MaterialVoxel.h"
#include "VX_Sim.h"
#include "VX_Voxel.h"
#include "XML_Rip.h"
#include "vx_mathtree.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

static std::string g_err;

struct vxo_sim {
    CVX_Sim* sim = nullptr;          // VXA path
    CVX_Environment* env = nullptr;
    CVX_Object* obj = nullptr;
    CVoxelyze* own = nullptr;        // desc path
    CVoxelyze* vx = nullptr;
    std::vector<int> matOf;          // voxel index -> material index
    std::map<CVX_Voxel*, int> vindex;
    std::vector<double> phase;
    std::vector<int> measured, fixedm;
    std::vector<double> thermalOn;
    // VX3 driver state
    double currentTime = 0.0, OptimalDt = 0.0, DtFrac = 0.9, dtFixed = 0.0;
    long CurStepCount = 0;
    bool vary = false, enableExpansion = false;
    double TempAmplitude = 0, TempPeriod = 0;
    Vec3D<double> initialCoM, currentCoM, comHist[2];
    int angleSampleTimes = 0;
    double recentAngle = 0;
    int state = VX3H_SIM_RUNNING;
    int maxSteps = 0;
    std::vector<vx3h_token> stopTok, fitTok;
    float lastDt = 0;
    VoxcadStop voxcad;               // VoxCAD stop-condition family and equilibrium mode
    int colInterval = 0;             // > 0: collision lists also rebuilt every colInterval steps
};

static int matIndex(vxo_sim* h, CVX_MaterialVoxel* m) {
    for (int i = 0; i < (int)h->vx->voxelMats.size(); i++) if (h->vx->voxelMats[i] == m) return i;
    return -1;
}

static void indexVoxels(vxo_sim* h) {
    h->vindex.clear();
    for (int i = 0; i < h->vx->voxelCount(); i++) h->vindex[h->vx->voxel(i)] = i;
}

static void updateCoM(vxo_sim* h) {
    double TotalMass = 0;
    Vec3D<double> Sum(0, 0, 0);
    for (int i = 0; i < h->vx->voxelCount(); i++) {
        CVX_Voxel* v = h->vx->voxel(i);
        if (!h->measured[matIndex(h, v->material())]) continue;
        double ThisMass = v->material()->mass();
        Sum += v->position() * ThisMass;
        TotalMass += ThisMass;
    }
    if (TotalMass == 0) { h->currentCoM = Vec3D<double>(); return; }
    h->currentCoM = Sum / TotalMass;
}

extern "C" {

const char* vxo_last_error(void) { return g_err.c_str(); }

vxo_sim* vxo_create(const vx3h_sim_desc* d) {
    if (!d || d->n_materials <= 0) { g_err = "empty description"; return nullptr; }
    vxo_sim* h = new vxo_sim;
    h->own = new CVoxelyze(0.001);
    CVoxelyze& Vx = *h->own;
    h->vx = h->own;
    Vx.setGravity(d->gravity_mult);
    Vx.enableFloor(d->floor_enabled != 0);
    for (int i = 0; i < d->n_materials; i++) {
        const vx3h_material_desc& m = d->materials[i];
        CVX_Material* p = Vx.addMaterial();
        p->matid = m.matid;
        switch (m.mat_model) {
        case VX3H_MDL_LINEAR: p->setModelLinear((float)m.elastic_mod); break;
        case VX3H_MDL_LINEAR_FAIL: p->setModelLinear((float)m.elastic_mod, (float)m.fail_stress); break;
        case VX3H_MDL_BILINEAR: p->setModelBilinear((float)m.elastic_mod, (float)m.plastic_mod, (float)m.yield_stress, (float)m.fail_stress); break;
        default: {
            std::vector<float> st, ss;
            for (int k = 0; k < m.n_data; k++) { ss.push_back((float)m.stress_data[k]); st.push_back((float)m.strain_data[k]); }
            p->setModel(m.n_data, st.data(), ss.data());
        }
        }
        p->setPoissonsRatio((float)m.poissons_ratio);
        p->setDensity((float)m.density);
        p->setCte((float)m.cte);
        p->setStaticFriction((float)m.u_static);
        p->setKineticFriction((float)m.u_dynamic);
        p->setInternalDamping((float)m.zeta_internal);
        p->setGlobalDamping((float)m.zeta_global);
        p->setCollisionDamping((float)m.zeta_collision);
        h->measured.push_back(m.is_measured);
        h->fixedm.push_back(m.fixed);
        h->thermalOn.push_back(m.thermal_on_after);
    }
    Vx.setVoxelSize(d->voxel_size);
    for (int i = 0; i < d->n_voxels; i++) {
        CVX_Voxel* v = Vx.setVoxel(Vx.material(d->material[i]), d->ijk[3 * i], d->ijk[3 * i + 1], d->ijk[3 * i + 2]);
        h->phase.push_back(d->phase ? d->phase[i] : 0.0);
        (void)v;
    }
    for (int e = 0; e < d->n_externals; e++) {
        const vx3h_external& x = d->externals[e];
        CVX_External* pe = Vx.voxel(x.voxel)->external();
        const dofComponent bits[6] = {X_TRANSLATE, Y_TRANSLATE, Z_TRANSLATE, X_ROTATE, Y_ROTATE, Z_ROTATE};
        for (int k = 0; k < 6; k++)
            if (x.dof_fixed & bits[k]) pe->setDisplacement(bits[k], k < 3 ? x.translate[k] : x.rotate[k - 3]);
        pe->addForce(Vec3D<float>((float)x.force[0], (float)x.force[1], (float)x.force[2]));
        pe->addMoment(Vec3D<float>((float)x.moment[0], (float)x.moment[1], (float)x.moment[2]));
    }
    if (d->temp_enabled) {
        float cur;
        if (d->vary_temp_enabled) {
            if (d->temp_period == 0) cur = 0.0f;
            else cur = (float)(d->temp_base + d->temp_amplitude * sin(2 * 3.1415926 / d->temp_period * 0.0));
        } else cur = (float)(d->temp_base + d->temp_amplitude);
        float t0 = (float)(cur - d->temp_base);
        for (int i = 0; i < Vx.voxelCount(); i++) Vx.voxel(i)->setTemperature(t0);
    }
    for (int i = 0; i < d->n_materials; i++) Vx.material(i)->setPoissonsRatio(d->volume_effects ? (float)d->materials[i].poissons_ratio : 0.0f);
    if (d->init_velocity)   // starting momentum (CVX_Voxel::linMom is a public member)
        for (int i = 0; i < Vx.voxelCount(); i++)
            Vx.voxel(i)->linMom = Vec3D<double>(d->init_velocity[3 * i], d->init_velocity[3 * i + 1], d->init_velocity[3 * i + 2]) *
                                  (double)Vx.voxel(i)->material()->mass();
    if (d->collisions_enabled) {
        Vx.boundingRadius = (float)d->bounding_radius;
        Vx.watchDistance = (float)d->watch_distance;
        Vx.enableCollisions(true);
    }
    h->DtFrac = d->dt_frac;
    h->dtFixed = d->dt_fixed;
    h->vary = d->vary_temp_enabled != 0;
    h->enableExpansion = d->enable_expansion != 0;
    h->TempAmplitude = d->temp_amplitude;
    h->TempPeriod = d->temp_period;
    h->maxSteps = d->max_steps;
    h->voxcad.type = d->stop_condition_type;
    h->voxcad.value = d->stop_condition_value;
    h->voxcad.eq = d->equilibrium_mode != 0;
    if (d->n_stop_tokens > 0) h->stopTok.assign(d->stop_tokens, d->stop_tokens + d->n_stop_tokens);
    if (d->n_fitness_tokens > 0) h->fitTok.assign(d->fitness_tokens, d->fitness_tokens + d->n_fitness_tokens);
    indexVoxels(h);
    updateCoM(h);
    h->initialCoM = h->currentCoM;
    return h;
}

/* Loads a VXA text through the reference's own reader and importer. */
vxo_sim* ref_load_vxa(const char* text) {
    vxo_sim* h = new vxo_sim;
    h->env = new CVX_Environment();
    h->obj = new CVX_Object();
    h->sim = new CVX_Sim();
    h->env->pObj = h->obj;
    h->sim->pEnv = h->env;
    CXML_Rip XML;
    std::string s(text);
    XML.fromXMLText(&s);
    std::string msg;
    h->sim->ReadVXA(&XML, &msg);
    h->sim->Import(NULL, NULL, &msg);
    h->vx = &h->sim->Vx;
    for (int i = 0; i < h->vx->materialCount(); i++) {
        CVX_Material* m = h->vx->material(i);
        h->measured.push_back(m->isMeasured);
        h->fixedm.push_back(m->fixed ? 1 : 0);
        h->thermalOn.push_back(m->TurnOnThermalExpansionAfterThisManySeconds);
    }
    for (int i = 0; i < h->vx->voxelCount(); i++) h->phase.push_back(h->vx->voxel(i)->phaseOffset);
    // VX3 driver settings the kernel constructor takes from the CPU model
    // (VX3_VoxelyzeKernel.cu:27-105): thermal variation, amplitude, period, DtFrac
    h->vary = h->env->IsTempVaryEnabled();
    h->TempAmplitude = h->env->GetTempAmplitude();
    h->TempPeriod = h->env->GetTempPeriod();
    h->DtFrac = h->sim->DtFrac;
    indexVoxels(h);
    return h;
}
/* Environment values the reference read (gravity multiplier, floor, temperatures of each voxel). */
void ref_vxa_env(vxo_sim* h, float* grav, int* floor, double* dtfrac, double* temp_amp, double* temp_period) {
    *grav = h->vx->gravity();
    *floor = h->vx->isFloorEnabled() ? 1 : 0;
    *dtfrac = h->sim ? h->sim->DtFrac : 0;
    *temp_amp = h->env ? h->env->GetTempAmplitude() : 0;
    *temp_period = h->env ? h->env->GetTempPeriod() : 0;
}
void ref_vxa_voxels(vxo_sim* h, int* ijk, int* mat, double* phase) {
    for (int i = 0; i < h->vx->voxelCount(); i++) {
        CVX_Voxel* v = h->vx->voxel(i);
        ijk[3 * i] = v->indexX(); ijk[3 * i + 1] = v->indexY(); ijk[3 * i + 2] = v->indexZ();
        mat[i] = matIndex(h, v->material());
        phase[i] = v->phaseOffset;
    }
}
/* Cilia data the reference's importer attached (VX_Sim.cpp:102-105, 381-386): per voxel
 * base and shift force (3 each), per material Cilia and TurnOnCiliaAfterThisManySeconds. */
void ref_vxa_cilia(vxo_sim* h, double* base, double* shift, double* matCilia, double* matOn) {
    for (int i = 0; i < h->vx->voxelCount(); i++) {
        CVX_Voxel* v = h->vx->voxel(i);
        for (int k = 0; k < 3; k++) { base[3 * i + k] = v->baseCiliaForce[k]; shift[3 * i + k] = v->shiftCiliaForce[k]; }
    }
    for (int i = 0; i < h->vx->materialCount(); i++) {
        matCilia[i] = h->vx->material(i)->Cilia;
        matOn[i] = h->vx->material(i)->TurnOnCiliaAfterThisManySeconds;
    }
}
/* Stop-condition type and value and the equilibrium-mode switch CVX_Sim::ReadVXA read
 * (VX_Sim.cpp:260-277). */
void ref_vxa_stop(vxo_sim* h, int* type, double* value, int* eq) {
    *type = h->sim ? (int)h->sim->GetStopConditionType() : 0;
    *value = h->sim ? h->sim->GetStopConditionValue() : 0;
    *eq = h->sim && h->sim->IsFeatureEnabled(VXSFEAT_EQUILIBRIUM_MODE) ? 1 : 0;
}
/* Colour channels [0,255] and matid of each material the importer made, in palette
 * order (CVX_Sim::CopyMat, VX_Sim.cpp:368-389); returns the material count. */
int ref_vxa_colors(vxo_sim* h, int* rgba, int* matid, int cap) {
    const int n = h->vx->materialCount();
    for (int i = 0; i < n && i < cap; i++) {
        CVX_Material* m = h->vx->material(i);
        rgba[4 * i] = m->red(); rgba[4 * i + 1] = m->green(); rgba[4 * i + 2] = m->blue(); rgba[4 * i + 3] = m->alpha();
        matid[i] = m->matid;
    }
    return n;
}
void ref_vxa_externals(vxo_sim* h, int* dof, double* trans, double* rot, float* force, float* moment) {
    for (int i = 0; i < h->vx->voxelCount(); i++) {
        CVX_Voxel* v = h->vx->voxel(i);
        dof[i] = 0;
        if (!v->externalExists()) continue;
        CVX_External* e = v->external();
        dof[i] = e->dofFixed;
        for (int k = 0; k < 3; k++) {
            trans[3 * i + k] = e->translation()[k]; rot[3 * i + k] = e->rotation()[k];
            force[3 * i + k] = e->force()[k]; moment[3 * i + k] = e->moment()[k];
        }
    }
}

void vxo_free(vxo_sim* h) {
    if (!h) return;
    delete h->own;
    delete h->sim;
    delete h->obj;
    delete h->env;
    delete h;
}
void vxo_counts(vxo_sim* h, int* nv, int* nl) {
    if (nv) *nv = h->vx->voxelCount();
    if (nl) *nl = h->vx->linkCount();
}
void vxo_topology(vxo_sim* h, int* vneg, int* vpos, int* axis, int* voxel_links) {
    std::map<CVX_Link*, int> li;
    for (int i = 0; i < h->vx->linkCount(); i++) {
        CVX_Link* l = h->vx->link(i);
        li[l] = i;
        if (vneg) vneg[i] = h->vindex[l->pVNeg];
        if (vpos) vpos[i] = h->vindex[l->pVPos];
        if (axis) axis[i] = (int)l->axis;
    }
    if (voxel_links)
        for (int v = 0; v < h->vx->voxelCount(); v++)
            for (int d = 0; d < 6; d++) {
                CVX_Link* l = h->vx->voxel(v)->links[d];
                voxel_links[6 * v + d] = l ? li[l] : -1;
            }
}
int vxo_do_time_step(vxo_sim* h, float dt) { return h->vx->doTimeStep(dt) ? 1 : 0; }
float vxo_recommended_dt(vxo_sim* h) { return h->vx->recommendedTimeStep(); }
void vxo_reset_time(vxo_sim* h) { h->vx->resetTime(); }
void vxo_set_gravity(vxo_sim* h, float g) { h->vx->setGravity(g); }
void vxo_enable_floor(vxo_sim* h, int en) { h->vx->enableFloor(en != 0); }
void vxo_enable_collisions(vxo_sim* h, int en) { h->vx->enableCollisions(en != 0); }
void vxo_set_ambient_temperature(vxo_sim* h, float t, int all) { h->vx->setAmbientTemperature(t, all != 0); }
void vxo_set_voxel_temperature(vxo_sim* h, int v, float t) { h->vx->voxel(v)->setTemperature(t); }
void vxo_ext_set_force(vxo_sim* h, int v, float fx, float fy, float fz) { h->vx->voxel(v)->external()->setForce(fx, fy, fz); }
void vxo_ext_set_moment(vxo_sim* h, int v, float mx, float my, float mz) { h->vx->voxel(v)->external()->setMoment(mx, my, mz); }
void vxo_ext_set_fixed(vxo_sim* h, int v, int m) { h->vx->voxel(v)->external()->setFixed(m & 1, m & 2, m & 4, m & 8, m & 16, m & 32); }
void vxo_ext_set_fixed_all(vxo_sim* h, int v, int fixed) { h->vx->voxel(v)->external()->setFixedAll(fixed != 0); }
void vxo_ext_set_displacement_all(vxo_sim* h, int v, double tx, double ty, double tz, double rx, double ry, double rz) {
    h->vx->voxel(v)->external()->setDisplacementAll(Vec3D<double>(tx, ty, tz), Vec3D<double>(rx, ry, rz));
}
void vxo_external_force(vxo_sim* h, int v, float* out) {
    Vec3D<float> r = h->vx->voxel(v)->externalForce();
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void vxo_material_set(vxo_sim* h, int m, int which, float value) {
    CVX_Material* p = h->vx->material(m);
    switch (which) {
    case 0: p->setInternalDamping(value); break;
    case 1: p->setGlobalDamping(value); break;
    case 2: p->setCollisionDamping(value); break;
    case 3: p->setStaticFriction(value); break;
    case 4: p->setKineticFriction(value); break;
    case 5: p->setCte(value); break;
    case 6: p->setPoissonsRatio(value); break;
    default: break;
    }
}
void vxo_replace_voxel_material(vxo_sim* h, int v, int m) {
    CVX_Voxel* pV = h->vx->voxel(v);
    h->vx->setVoxel(h->vx->material(m), pV->indexX(), pV->indexY(), pV->indexZ());
}

static double recDtVX3(vxo_sim* h) {
    double MaxFreq2 = 0.0f;
    for (int i = 0; i < h->vx->linkCount(); i++) {
        CVX_Link* pL = h->vx->link(i);
        double m1 = pL->pVNeg->mat->mass(), m2 = pL->pVPos->mat->mass();
        double thisMaxFreq2 = pL->axialStiffness() / (m1 < m2 ? m1 : m2);
        if (thisMaxFreq2 > MaxFreq2) MaxFreq2 = thisMaxFreq2;
    }
    if (MaxFreq2 <= 0.0f) {
        for (int i = 0; i < h->vx->voxelCount(); i++) {
            CVX_Voxel* v = h->vx->voxel(i);
            double t = v->mat->youngsModulus() * v->mat->nomSize / v->mat->mass();
            if (t > MaxFreq2) MaxFreq2 = t;
        }
    }
    if (MaxFreq2 <= 0.0f) return 0.0f;
    return 1.0f / (6.283185f * sqrt(MaxFreq2));
}

static bool stopMet(vxo_sim* h) {
    if (h->maxSteps > 0 && h->CurStepCount >= h->maxSteps) return true;
    if (!h->stopTok.empty() &&
        vxo_mathtree_eval(h->stopTok.data(), (int)h->stopTok.size(), h->currentCoM.x, h->currentCoM.y, h->currentCoM.z, 0, h->currentTime,
                          h->recentAngle, 0, 0, h->vx->voxelCount()) > 0)
        return true;
    return h->voxcad.type != 0 && h->voxcad.met(h->CurStepCount, h->currentTime, h->TempPeriod, h->vx->voxelCount());
}

static bool stepVX3(vxo_sim* h) {
    if (h->vary && h->TempPeriod > 0) {
        for (int i = 0; i < h->vx->voxelCount(); i++) {
            CVX_Voxel* v = h->vx->voxel(i);
            int mi = matIndex(h, v->material());
            if (h->thermalOn[mi] > h->currentTime) continue;
            if (h->fixedm[mi]) continue;
            double T = h->TempAmplitude * sin(2 * 3.1415926f * (h->currentTime / h->TempPeriod + h->phase[i]));
            if (!h->enableExpansion) { if (T > 0) T = 0; }
            v->setTemperature((float)T);
        }
    }
    h->CurStepCount++;
    float dt;
    if (h->dtFixed > 0) dt = (float)h->dtFixed;
    else {
        if (!h->OptimalDt) h->OptimalDt = recDtVX3(h);
        if (h->OptimalDt < 1e-10) h->OptimalDt = 1e-10;
        dt = h->DtFrac * h->OptimalDt;
    }
    h->lastDt = dt;
    // an optional fixed broadphase period on top of CVoxelyze's staleness rule
    if (h->colInterval > 0 && (h->CurStepCount - 1) % h->colInterval == 0) h->vx->collisionsStale = true;
    if (!h->vx->doTimeStep(dt)) return false;
    int CycleStep = int(h->TempPeriod / dt);
    if (CycleStep > 0 && h->CurStepCount % CycleStep == 0) {
        h->angleSampleTimes++;
        h->comHist[0] = h->comHist[1];
        h->comHist[1] = h->currentCoM;
        updateCoM(h);
        Vec3D<double> A = h->comHist[0], B = h->comHist[1], C = h->currentCoM;
        if (B == C || A == B || h->angleSampleTimes < 3) h->recentAngle = 0;
        else h->recentAngle = acos((B - A).Dot(C - B) / (B.Dist(A) * C.Dist(B)));
    }
    h->currentTime += dt;
    CVoxelyze* vx = h->vx;
    h->voxcad.afterStep(
        vx->isCollisionsEnabled(), h->dtFixed > 0 ? h->dtFixed : h->DtFrac * h->OptimalDt,
        [&] { for (int i = 0; i < vx->voxelCount(); i++) vx->voxel(i)->haltMotion(); },
        [&](bool vel, bool kin, bool strainE, double& maxVel2, double& totKE, double& totSE) {
            for (int i = 0; i < vx->voxelCount(); i++) {
                CVX_Voxel* v = vx->voxel(i);
                if (vel) { const double v2 = v->velocity().Length2(); if (v2 > maxVel2) maxVel2 = v2; }
                if (kin) totKE += v->kineticEnergy();
            }
            if (strainE) for (int i = 0; i < vx->linkCount(); i++) totSE += vx->link(i)->strainEnergy();
        });
    return true;
}

int vxo_vx3_step(vxo_sim* h, int n) {
    int done = 0;
    for (int j = 0; j < n; j++) {
        if (h->state != VX3H_SIM_RUNNING) break;
        if (stopMet(h)) { h->state = VX3H_SIM_DONE; break; }
        if (!stepVX3(h)) { h->state = VX3H_SIM_DIVERGED; break; }
        done++;
    }
    if (h->state == VX3H_SIM_RUNNING && stopMet(h)) h->state = VX3H_SIM_DONE;
    return done;
}
void vxo_vx3_status(vxo_sim* h, vx3h_sim_status* st) {
    Vec3D<double> save = h->currentCoM;
    updateCoM(h);
    Vec3D<double> off = h->currentCoM - h->initialCoM;
    memset(st, 0, sizeof(*st));
    st->state = h->state;
    st->steps = (int)h->CurStepCount;
    st->num_voxel = h->vx->voxelCount();
    st->current_time = h->currentTime;
    st->dt = h->lastDt;
    st->recommended_dt = h->OptimalDt;
    st->initial_com[0] = h->initialCoM.x; st->initial_com[1] = h->initialCoM.y; st->initial_com[2] = h->initialCoM.z;
    st->current_com[0] = h->currentCoM.x; st->current_com[1] = h->currentCoM.y; st->current_com[2] = h->currentCoM.z;
    st->fitness = vxo_mathtree_eval(h->fitTok.data(), (int)h->fitTok.size(), off.x, off.y, off.z, 0, h->currentTime, h->recentAngle, 0, 0,
                                    h->vx->voxelCount());
    st->recent_angle = h->recentAngle;
    int nm = 0;
    double tot = 0;
    for (int i = 0; i < h->vx->voxelCount(); i++) {
        CVX_Voxel* v = h->vx->voxel(i);
        if (!h->measured[matIndex(h, v->material())]) continue;
        nm++;
        tot += v->position().Dist(v->originalPosition());
    }
    st->num_measured_voxel = nm;
    st->total_distance_of_all_voxels = tot;
    h->currentCoM = save;
}
void vxo_voxel_state(vxo_sim* h, double* pos, double* orient, double* lm, double* am, int* flags, float* temp) {
    for (int i = 0; i < h->vx->voxelCount(); i++) {
        CVX_Voxel* v = h->vx->voxel(i);
        if (pos) { pos[3 * i] = v->pos.x; pos[3 * i + 1] = v->pos.y; pos[3 * i + 2] = v->pos.z; }
        if (orient) { orient[4 * i] = v->orient.w; orient[4 * i + 1] = v->orient.x; orient[4 * i + 2] = v->orient.y; orient[4 * i + 3] = v->orient.z; }
        if (lm) { lm[3 * i] = v->linMom.x; lm[3 * i + 1] = v->linMom.y; lm[3 * i + 2] = v->linMom.z; }
        if (am) { am[3 * i] = v->angMom.x; am[3 * i + 1] = v->angMom.y; am[3 * i + 2] = v->angMom.z; }
        if (flags) flags[i] = v->boolStates;
        if (temp) temp[i] = v->temp;
    }
}
void vxo_link_state(vxo_sim* h, float* strain, float* maxStrain, float* off, int* sa, int* bs, double* rl, double* f12) {
    for (int i = 0; i < h->vx->linkCount(); i++) {
        CVX_Link* l = h->vx->link(i);
        if (strain) strain[i] = l->strain;
        if (maxStrain) maxStrain[i] = l->maxStrain;
        if (off) off[i] = l->strainOffset;
        if (sa) sa[i] = l->smallAngle ? 1 : 0;
        if (bs) bs[i] = l->boolStates;
        if (rl) rl[i] = l->currentRestLength;
        if (f12) {
            const Vec3D<double>* vs[4] = {&l->forceNeg, &l->forcePos, &l->momentNeg, &l->momentPos};
            for (int k = 0; k < 4; k++) { f12[12 * i + 3 * k] = vs[k]->x; f12[12 * i + 3 * k + 1] = vs[k]->y; f12[12 * i + 3 * k + 2] = vs[k]->z; }
        }
    }
}
void vxo_local_signal(vxo_sim* h, double* out) {   // the CPU engine has no signals
    for (int i = 0; i < h->vx->voxelCount(); i++) out[i] = 0;
}
void vxo_voxcad_history(vxo_sim* h, double* k, double* t, double* m) {
    const VoxcadStop& c = h->voxcad;
    for (int i = 0; i < VoxcadStop::HIST_LEN; i++) { k[i] = c.KinEHistory[i]; t[i] = c.TotEHistory[i]; m[i] = c.MaxMoveHistory[i]; }
}
void vxo_corner_offset(vxo_sim* h, int v, int corner, float* out) {
    Vec3D<float> r = h->vx->voxel(v)->cornerOffset((CVX_Voxel::voxelCorner)corner);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void vxo_set_collision_interval(vxo_sim* h, int n) { h->colInterval = n > 0 ? n : 0; }
int vxo_collision_pairs(vxo_sim* h, int* pairs, int cap) {
    const std::vector<CVX_Collision*>* cl = h->vx->collisionList();
    int n = (int)cl->size();
    for (int i = 0; i < n && i < cap; i++) { pairs[2 * i] = h->vindex[(*cl)[i]->voxel1()]; pairs[2 * i + 1] = h->vindex[(*cl)[i]->voxel2()]; }
    return n;
}
int vxo_voxel_material_consts(vxo_sim* h, float* out, int cap) {
    int n = (int)h->vx->voxelMats.size();
    for (int i = 0; i < n && i < cap; i++) {
        CVX_MaterialVoxel* m = h->vx->voxelMats[i];
        float* o = out + 14 * i;
        o[0] = m->_mass; o[1] = m->_massInverse; o[2] = m->_sqrtMass; o[3] = m->_momentInertia; o[4] = m->_momentInertiaInverse;
        o[5] = m->_2xSqMxExS; o[6] = m->_2xSqIxExSxSxS; o[7] = m->E; o[8] = m->nu; o[9] = m->_eHat; o[10] = m->penetrationStiffness();
        o[11] = m->gravityForce(); o[12] = m->muStatic; o[13] = m->muKinetic;
    }
    return n;
}
int vxo_link_material_consts(vxo_sim* h, int* link_mat_index, float* out, int cap) {
    std::vector<CVX_MaterialLink*> mats(h->vx->linkMats.begin(), h->vx->linkMats.end());
    if (link_mat_index)
        for (int i = 0; i < h->vx->linkCount(); i++)
            link_mat_index[i] = (int)(std::find(mats.begin(), mats.end(), h->vx->link(i)->mat) - mats.begin());
    int n = (int)mats.size();
    for (int i = 0; i < n && i < cap; i++) {
        CVX_MaterialLink* m = mats[i];
        float* o = out + 20 * i;
        o[0] = m->_a1; o[1] = m->_a2; o[2] = m->_b1; o[3] = m->_b2; o[4] = m->_b3; o[5] = m->_sqA1; o[6] = m->_sqA2xIp; o[7] = m->_sqB1;
        o[8] = m->_sqB2xFMp; o[9] = m->_sqB3xIp; o[10] = m->E; o[11] = m->nu; o[12] = m->_eHat; o[13] = m->linear ? 1.f : 0.f;
        o[14] = m->epsilonFail; o[15] = m->epsilonYield; o[16] = m->alphaCTE; o[17] = (float)m->strainData.size();
        o[18] = m->strainData.size() > 1 ? m->strainData[1] : 0; o[19] = m->stressData.size() > 1 ? m->stressData[1] : 0;
    }
    return n;
}

/* The reference's MathTree evaluator is a CUDA header (Utils/VX3_MathTree.h); its
 * algorithm is restated in vx_mathtree.h. */
double vxo_mathtree_eval(const vx3h_token* toks, int n, double x, double y, double z, double hit, double t, double angle, double closeness,
                         int ncp, int nv) {
    return vxo::mathtree_eval(toks, n, x, y, z, hit, t, angle, closeness, ncp, nv);
}

}  // extern "C"
