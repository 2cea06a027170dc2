// sanitize_main.cpp -- drives the CPU oracle (test infrastructure only) under the
// address/undefined-behaviour and thread sanitizers.  Four simulations that between
// them reach every part of the restatement with dynamic memory -- collision lists and
// their regeneration, nearby lists rebuilt after topology changes (detach, attach),
// target closeness, Poisson strains -- run to completion, each in its own thread (the
// tests call the oracle from several threads at once, one simulation per thread).
//
// Built and run by `make -C oracle sanitize`; a leak, an out-of-bounds access, an
// undefined operation or a data race ends the program with a non-zero status.
#include <cstdio>
#include <thread>
#include <vector>

#include "vx_oracle.h"

namespace {

vx3h_material_desc material(double cte, double fail, int sticky, int target) {
    vx3h_material_desc m = {};
    m.mat_model = fail > 0 ? VX3H_MDL_LINEAR_FAIL : VX3H_MDL_LINEAR;
    m.elastic_mod = 1e6;
    m.fail_stress = fail;
    m.density = 1000;
    m.poissons_ratio = 0.35;
    m.cte = cte;
    m.u_static = 1;
    m.u_dynamic = 0.5;
    m.zeta_internal = 1;
    m.zeta_global = 0.01;
    m.zeta_collision = 0.8;
    m.is_measured = 1;
    m.sticky = sticky;
    m.is_target = target;
    return m;
}

struct Case {
    const char* name;
    std::vector<vx3h_material_desc> mats;
    std::vector<int> ijk, mat;
    vx3h_sim_desc d = {};
    int steps = 0;
};

void finish(Case& c) {
    c.d.voxel_size = 0.01;
    c.d.n_materials = (int)c.mats.size();
    c.d.materials = c.mats.data();
    c.d.n_voxels = (int)c.mat.size();
    c.d.ijk = c.ijk.data();
    c.d.material = c.mat.data();
    c.d.gravity_mult = 1;
    c.d.floor_enabled = 1;
    c.d.dt_frac = 0.9;
    c.d.volume_effects = 1;
    c.d.collisions_enabled = 1;
    c.d.bounding_radius = 0.75;
    c.d.watch_distance = 1.0;
    c.d.temp_enabled = c.d.vary_temp_enabled = 1;
    c.d.temp_amplitude = 10;
    c.d.temp_period = 0.02;
    c.d.max_steps = c.steps;
}

void add(Case& c, int x, int y, int z, int m) {
    c.ijk.insert(c.ijk.end(), {x, y, z});
    c.mat.push_back(m);
}

// 3x3x3 cube whose centre voxel contracts until its links fail (detach), with collisions
Case shrinkingCore() {
    Case c;
    c.name = "detach";
    c.mats = {material(0.0, 1e4, 0, 0), material(0.02, 1e4, 0, 0)};
    for (int z = 0; z < 3; z++)
        for (int y = 0; y < 3; y++)
            for (int x = 0; x < 3; x++) add(c, x, y, z, (x == 1 && y == 1 && z == 1) ? 1 : 0);
    c.steps = 1500;
    finish(c);
    c.d.enable_detach = 1;
    return c;
}

// two sticky bars, one dropped across the other (attach, with collisions)
Case stickyBars() {
    Case c;
    c.name = "attach";
    c.mats = {material(0.0, 0, 1, 0)};
    for (int x = 0; x < 6; x++) add(c, x, 2, 0, 0);
    for (int y = 0; y < 6; y++) add(c, 2, y, 2, 0);
    c.steps = 1500;
    finish(c);
    c.d.enable_attach = 1;
    return c;
}

// a 4x4x2 slab with a target material on top: target pairs and closeness
Case targets() {
    Case c;
    c.name = "targets";
    c.mats = {material(0.01, 0, 0, 0), material(0.0, 0, 0, 1)};
    for (int z = 0; z < 2; z++)
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) add(c, x, y, z, z == 1 && (x + y) % 2 ? 1 : 0);
    c.steps = 1500;
    finish(c);
    c.d.max_dist_close_pairs = 1.5;
    c.d.temp_period = 0.005;
    return c;
}

// 5x5x5 cube falling on its edge: many collision regenerations
Case fallingCube() {
    Case c;
    c.name = "cube";
    c.mats = {material(0.01, 0, 0, 0)};
    for (int z = 0; z < 5; z++)
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++) add(c, x, y, z + 2, 0);
    c.steps = 800;
    finish(c);
    return c;
}

int run(Case* c) {
    vxo_sim* s = vxo_create(&c->d);
    if (!s) {
        std::fprintf(stderr, "%s: create failed: %s\n", c->name, vxo_last_error());
        return 1;
    }
    const int done = vxo_vx3_step(s, c->steps);
    vx3h_sim_status st;
    vxo_vx3_status(s, &st);
    std::vector<int> pairs(4096);
    const int np = vxo_collision_pairs(s, pairs.data(), (int)pairs.size() / 2);
    std::printf("%s: %d steps, state %d, %d collision pairs, %d close pairs\n", c->name, done, st.state, np,
                st.num_close_pairs);
    vxo_free(s);
    return done == c->steps ? 0 : 1;
}

}  // namespace

int main() {
    std::vector<Case> cases = {shrinkingCore(), stickyBars(), targets(), fallingCube()};
    for (Case& c : cases) {   // descriptors point into the cases' own vectors
        c.d.materials = c.mats.data();
        c.d.ijk = c.ijk.data();
        c.d.material = c.mat.data();
    }
    std::vector<int> rc(cases.size(), 0);
    std::vector<std::thread> th;
    for (size_t i = 0; i < cases.size(); i++) th.emplace_back([&, i] { rc[i] = run(&cases[i]); });
    for (auto& t : th) t.join();
    int bad = 0;
    for (int r : rc) bad |= r;
    return bad;
}
