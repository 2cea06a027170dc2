"""Known-answer tests that pin the CPU oracle (oracle/vx_oracle.cpp) to the
reference's own unit tests and to closed forms:

* VoxCAD/src/VX2/test/tVoxelyze.h:142-198   singleBondFixedFree: F/k and torsion/bending
* VoxCAD/src/VX2/test/tVoxelyze.h:415-443   freqency: axial period 2*pi/sqrt(k/m)
* VoxCAD/src/VX2/test/tVoxelyze.h:1034-1114 staticFriction: no sliding below mu_s*m*g
* VoxCAD/src/VX2/test/tVX_MaterialLink.h:3-37  harmonic-mean link modulus 20/11
* VoxCAD/src/VX2/test/tVX_MaterialVoxel.h:4-12 mass and moment of inertia of a voxel
* thermal expansion of a free bar: length change n*s*CTE*T
"""
import numpy as np
import pytest

from gpuvoxels_amd.model import MT, Ext, Material, Model
from oracle_lib import OSim

TX, TY, TZ = 1, 2, 4
RX, RY, RZ = 8, 16, 32
ALL = 63
OFFSETS = {"X_POS": (1, 0, 0), "X_NEG": (-1, 0, 0), "Y_POS": (0, 1, 0), "Y_NEG": (0, -1, 0),
           "Z_POS": (0, 0, 1), "Z_NEG": (0, 0, -1)}


def two_voxels(direction, free_dofs_fixed, force=(0, 0, 0), moment=(0, 0, 0)):
    """test2Vox of tVoxelyze.h:64-115: voxel 1 fixed at the origin, voxel 2 beside it in
    `direction`, loaded and constrained in all dofs but one; E = 1e6, rho = 1e3,
    s = 1 mm, internal damping 1, global damping 0.2."""
    off = OFFSETS[direction]
    m = Model(voxel_size=0.001)
    m.materials = [Material(elastic_mod=1e6, density=1e3, zeta_internal=1.0, zeta_global=0.2)]
    cells = [(0, 0, 0), off]
    shift = [-min(0, off[k]) for k in range(3)]   # lattice indices must be non-negative
    ijk = np.array([[c[k] + shift[k] for k in range(3)] for c in cells], np.int32)
    order = np.lexsort((ijk[:, 0], ijk[:, 1], ijk[:, 2]))
    m.ijk = ijk[order]
    m.mat = np.zeros(2, np.int32)
    first = int(np.where(order == 0)[0][0])
    second = 1 - first
    m.externals = [Ext(first, dof_fixed=ALL), Ext(second, dof_fixed=free_dofs_fixed, force=force, moment=moment)]
    return m, second, off


def settle(m, second, off, what, expected, max_steps=1000):
    """Number of steps until the watched quantity stays within 1e-5 relative of the
    expected value for 10 consecutive steps (the test2Vox convergence rule)."""
    o = OSim(m)
    ts = o.recommended_dt()
    run = 0
    for k in range(max_steps):
        o.step_raw(ts)
        st = o.state()
        if what < 3:
            ref = m.ijk[second] * m.voxel_size
            data = st["pos"][second][what] - ref[what]
        else:
            q = st["orient"][second]
            # rotation vector of the (w, x, y, z) quaternion, small angles
            ang = 2 * np.arccos(np.clip(q[0], -1, 1))
            s = np.sqrt(max(1 - q[0] * q[0], 0.0))
            rv = np.zeros(3) if s < 1e-12 else ang * q[1:] / s
            data = rv[what - 3]
        if abs(data - expected) < abs(expected) * 1e-5 or np.float32(data) == np.float32(expected):
            run += 1
        else:
            run = 0
        if run == 10:
            return k
    return max_steps


@pytest.mark.parametrize("direction,force,free,axis,expected,limit", [
    ("X_POS", (1e-3, 0, 0), ALL & ~TX, 0, 1e-6, 150),
    ("X_POS", (-1e-3, 0, 0), ALL & ~TX, 0, -1e-6, 150),
    ("X_POS", (0, 1e-3, 0), ALL & ~TY, 1, 1e-6, 150),
    ("X_POS", (0, 0, -1e-3), ALL & ~TZ, 2, -1e-6, 150),
    ("Y_NEG", (0, 1e-3, 0), ALL & ~TY, 1, 1e-6, 150),
    ("Z_POS", (1e-3, 0, 0), ALL & ~TX, 0, 1e-6, 150),
])
def test_single_bond_force(direction, force, free, axis, expected, limit):
    m, second, off = two_voxels(direction, free, force=force)
    assert settle(m, second, off, axis, expected) < limit


@pytest.mark.parametrize("direction,moment,free,axis,expected,limit", [
    ("X_POS", (1e-9, 0, 0), ALL & ~RX, 3, 1.2e-5, 200),    # torsion
    ("X_POS", (0, 1e-9, 0), ALL & ~RY, 4, 3e-6, 200),      # bending
    ("Y_POS", (0, 1e-9, 0), ALL & ~RY, 4, 1.2e-5, 200),
    ("Z_POS", (1e-9, 0, 0), ALL & ~RX, 3, 3e-6, 200),
])
def test_single_bond_moment(direction, moment, free, axis, expected, limit):
    m, second, off = two_voxels(direction, free, moment=moment)
    assert settle(m, second, off, axis, expected) < limit


def calc_period(ts, data):   # zero crossings, tVoxelyze.h:43-62
    zc = []
    for i in range(1, len(data)):
        if (data[i - 1] <= 0 < data[i]) or (data[i - 1] >= 0 > data[i]):
            perc = data[i - 1] / (data[i - 1] - data[i])
            zc.append(ts * (i - 1 + perc))
    return np.mean(np.diff(zc)) * 2


def test_axial_frequency():
    m = Model(voxel_size=0.001)
    m.materials = [Material(elastic_mod=1e6, density=1e3, zeta_internal=0.0, zeta_global=0.0)]
    m.ijk = np.array([[0, 0, 0], [1, 0, 0]], np.int32)
    m.mat = np.zeros(2, np.int32)
    m.externals = [Ext(0, dof_fixed=ALL), Ext(1, dof_fixed=ALL & ~TX, force=(1e-3, 0, 0))]
    o = OSim(m)
    ts = o.recommended_dt() / 10
    data = []
    for _ in range(1000):
        o.step_raw(ts)
        data.append(o.state()["pos"][1][0] - 0.001001)
    expected = 2 * 3.1415926 / np.sqrt(1e9)
    assert calc_period(ts, data) == pytest.approx(expected, abs=expected / 1000)


def test_static_friction_holds_below_threshold():
    vs, rho = 0.001, 1e3
    m = Model(voxel_size=vs)
    m.materials = [Material(elastic_mod=1e6, density=rho, u_static=1.0, u_dynamic=0.1, zeta_global=1.0)]
    m.ijk = np.array([[0, 0, 0]], np.int32)
    m.mat = np.zeros(1, np.int32)
    m.floor_enabled = 1
    m.gravity_mult = 1.0
    o = OSim(m)
    ts = o.recommended_dt()
    for _ in range(50):
        o.step_raw(ts)
    normal = np.float32(rho * vs ** 3 * 9.80665)
    x0 = o.state()["pos"][0][0]
    o.set_force(0, (0.9 * normal, 0, 0))     # just below mu_s * m * g
    for _ in range(200):
        o.step_raw(ts)
    assert o.state()["pos"][0][0] == x0
    o.set_force(0, (1.1 * normal, 0, 0))     # above: it slides
    for _ in range(200):
        o.step_raw(ts)
    assert o.state()["pos"][0][0] > x0


def test_link_material_harmonic_mean():
    m = Model(voxel_size=0.001)
    m.materials = [Material(elastic_mod=1e6, density=1e3), Material(elastic_mod=1e7, density=1e3)]
    m.ijk = np.array([[0, 0, 0], [1, 0, 0]], np.int32)
    m.mat = np.array([0, 1], np.int32)
    idx, consts = OSim(m).link_material_consts()
    assert consts[idx[0], 10] == pytest.approx(20.0 / 11.0 * 1e6, rel=1e-6)   # E of the link


def test_voxel_material_mass_and_inertia():
    s, rho = 0.002, 100.0   # tVX_MaterialVoxel.h: mass 8e-7, I = mass*s^2/6
    m = Model(voxel_size=s)
    m.materials = [Material(elastic_mod=1e6, density=rho)]
    m.ijk = np.array([[0, 0, 0]], np.int32)
    m.mat = np.zeros(1, np.int32)
    c = OSim(m).voxel_material_consts()[0]
    assert c[0] == pytest.approx(8e-7, rel=1e-6)
    assert c[3] == pytest.approx(8e-7 * s * s / 6, rel=1e-6)


def test_free_bar_thermal_expansion():
    n, s, cte, temp = 6, 0.01, 0.01, 5.0
    m = Model(voxel_size=s)
    m.materials = [Material(elastic_mod=1e6, density=1e3, cte=cte, zeta_internal=1.0, zeta_global=0.5)]
    m.ijk = np.array([[i, 0, 0] for i in range(n)], np.int32)
    m.mat = np.zeros(n, np.int32)
    m.temp_enabled = 1
    m.temp_base = 25.0
    m.temp_amplitude = temp     # no variation: every voxel sits at TempBase + A, i.e. T = A
    o = OSim(m)
    for _ in range(20000):
        if o.step(1) == 0:
            break
    p = o.state()["pos"]
    # each of the n-1 links rests at s*(1 + CTE*T)
    assert p[-1][0] - p[0][0] == pytest.approx((n - 1) * s * (1 + cte * temp), rel=1e-5)


# ---- VX3 driver semantics (Voxelyze3/src/VX3): closed forms and invariants

def row_of_voxels(phases, expansion=0, amplitude=10.0, period=0.1):
    m = Model(voxel_size=0.01)
    # E = 1e3 keeps the fixed 1 ms step below the stable 1.6 ms
    m.materials = [Material(elastic_mod=1e3, density=1e3, cte=0.01)]
    m.ijk = np.array([[i, 0, 0] for i in range(len(phases))], np.int32)
    m.mat = np.zeros(len(phases), np.int32)
    m.phase = np.array(phases, np.float64)
    m.temp_enabled = m.vary_temp_enabled = 1
    m.temp_amplitude, m.temp_period = amplitude, period
    m.enable_expansion = expansion
    m.dt_fixed = 0.001
    return m


@pytest.mark.parametrize("expansion", [0, 1])
def test_phase_temperature_closed_form(expansion):
    """T_v = A sin(2 pi (t/T_p + phi_v)) with phi in cycles, clamped to <= 0 unless
    expansion is enabled (VX3_VoxelyzeKernel.cu:625-650): at t = 0 the phases
    0, 1/4, 1/2, 3/4 give 0, A, 0, -A."""
    a = 10.0
    o = OSim(row_of_voxels([0.0, 0.25, 0.5, 0.75], expansion, a))
    o.step(1)                       # temperature is set from t = 0 before the first step
    t = o.state()["temp"]
    want = np.array([0.0, a if expansion else 0.0, 0.0, -a])
    assert np.allclose(t, want, atol=1e-5 * a)


def test_phase_is_a_time_shift():
    """A voxel at phase 1/4 and t = 0 has the temperature a voxel at phase 0 has at
    t = T_p/4 (25 steps of 1 ms with T_p = 0.1 s)."""
    o = OSim(row_of_voxels([0.0, 0.25], 1))
    o.step(1)
    t_shift = o.state()["temp"][1]
    o.step(25)
    assert o.state()["temp"][0] == pytest.approx(t_shift, rel=1e-4)


def test_recommended_dt_closed_form():
    """dt = DtFrac / (2 pi sqrt(k/m)), k = E s for a nu = 0 link of length s,
    m = rho s^3 (Voxelyze.cpp:286-311, VX3_VoxelyzeKernel.cu:184-217,245-253)."""
    s, e, rho = 0.01, 1e6, 1e3
    m = Model(voxel_size=s)
    m.materials = [Material(elastic_mod=e, density=rho)]
    m.ijk = np.array([[0, 0, 0], [1, 0, 0]], np.int32)
    m.mat = np.zeros(2, np.int32)
    m.dt_frac = 0.9
    o = OSim(m)
    o.step(1)
    want = 0.9 / (2 * np.pi * np.sqrt(e * s / (rho * s ** 3)))
    assert o.status().dt == pytest.approx(want, rel=1e-6)


def test_center_of_mass_measured_voxels_only():
    """The CoM is mass weighted over measured voxels (VX3_VoxelyzeKernel.cu:477-493):
    a symmetric 4^3 cube centres at 1.5 s; an unmeasured extra column is ignored."""
    s = 0.01
    m = Model(voxel_size=s)
    m.materials = [Material(elastic_mod=1e6, density=1e3), Material(elastic_mod=1e6, density=5e3, is_measured=0)]
    cube = [(i, j, k) for k in range(4) for j in range(4) for i in range(4)]
    extra = [(4, 0, k) for k in range(4)]
    m.ijk = np.array(cube + extra, np.int32)
    m.mat = np.array([0] * len(cube) + [1] * len(extra), np.int32)
    o = OSim(m)
    o.step(1)
    assert np.allclose(list(o.status().initial_com), [1.5 * s] * 3, rtol=0, atol=1e-12)


def test_fitness_and_stop_condition():
    """Fitness is the MathTree of the CoM displacement (here its 3D length); the
    stop formula t > 0.01 ends the run at the first step count n with n*dt > 0.01,
    dt being the float-rounded 1 ms (VX3_SimulationManager.cu:62-115)."""
    from gpuvoxels_amd.model import FITNESS_DISTANCE_3D, MT
    m = row_of_voxels([0.0, 0.5, 0.25], 0)
    m.fitness_tokens = FITNESS_DISTANCE_3D
    m.stop_tokens = [(MT["mtCONST"], 0.01), (MT["mtVAR"], 4.0), (MT["mtGREATERTHAN"], 0.0), (MT["mtEND"], 0.0)]
    m.gravity_mult = 1.0
    m.floor_enabled = 1
    o = OSim(m)
    o.step(1000)
    st = o.status()
    dt = float(np.float32(0.001))
    n = next(k for k in range(1, 100) if k * dt > 0.01)
    assert st.state == 1 and st.steps == n
    d = np.array(list(st.current_com)) - np.array(list(st.initial_com))
    assert d[2] < -1e-4       # it fell onto the floor
    assert st.fitness == pytest.approx(np.linalg.norm(d), rel=1e-12)


def test_corner_offsets_split_strain_between_halves():
    """CVX_Voxel::cornerOffset (VX_Voxel.cpp / VX3_Voxel.cu:147-159), the NNN and PPP
    corners the history stream prints: a stretched link between a soft (E1) and a stiff
    (E2) voxel splits its strain as two springs in series, each half straining in
    inverse proportion to its modulus (2*s*E_other/(E1+E2)); faces without a link stay
    at half the voxel size; the two halves together extend by the whole link."""
    size, e1, e2 = 0.001, 1e6, 4e6
    m = Model(voxel_size=size)
    m.materials = [Material(elastic_mod=e1, density=1e3, zeta_internal=1.0, zeta_global=0.2),
                   Material(elastic_mod=e2, density=1e3, zeta_internal=1.0, zeta_global=0.2)]
    m.ijk = np.array([[0, 0, 0], [1, 0, 0]], np.int32)
    m.mat = np.array([0, 1], np.int32)
    m.externals = [Ext(0, dof_fixed=ALL), Ext(1, dof_fixed=ALL & ~TX, force=(0.5, 0, 0))]
    o = OSim(m)
    ts = o.recommended_dt()
    for _ in range(300):
        o.step_raw(ts)
    s = float(o.link_state()["strain"][0])
    assert s > 1e-5
    h = 0.5 * size
    n0, p0 = o.corner_offset(0, 0), o.corner_offset(0, 7)
    n1, p1 = o.corner_offset(1, 0), o.corner_offset(1, 7)
    half0 = 2 * s * e2 / (e1 + e2)      # the soft voxel's half
    half1 = 2 * s * e1 / (e1 + e2)
    assert p0[0] == pytest.approx(h * (1 + half0), rel=1e-6)
    assert n1[0] == pytest.approx(-h * (1 + half1), rel=1e-6)
    assert n0[0] == np.float32(-h) and p1[0] == np.float32(h)
    for c in (n0, n1):
        assert c[1] == np.float32(-h) and c[2] == np.float32(-h)
    for c in (p0, p1):
        assert c[1] == np.float32(h) and c[2] == np.float32(h)
    assert (p0[0] - h) + (-n1[0] - h) == pytest.approx(size * s, rel=1e-5)


def test_corner_offsets_follow_thermal_size():
    """Without links the corners sit at half the thermal size, 0.5*s*(1 + T*CTE)."""
    m = Model(voxel_size=0.01)
    m.materials = [Material(elastic_mod=1e6, density=1e3, cte=0.01)]
    m.ijk = np.array([[0, 0, 0]], np.int32)
    m.mat = np.zeros(1, np.int32)
    o = OSim(m)
    o.lib.vxo_set_voxel_temperature(o.h, 0, np.float32(12.5).item())
    h = 0.5 * 0.01 * float(np.float32(1) + np.float32(12.5) * np.float32(0.01))
    assert np.array_equal(o.corner_offset(0, 7), np.float32([h, h, h]))
    assert np.array_equal(o.corner_offset(0, 0), np.float32([-h, -h, -h]))


# ---- VoxCAD stop-condition family and equilibrium mode (VoxCAD/src/VX1/VX_Sim.cpp:862-891,
# 1015-1160, 1833-1837), on top of the VX3 step

def _drop_cube(sc=0, value=0.0, eq=0, lift=2, n=3):
    m = Model(voxel_size=0.01)
    m.materials = [Material(elastic_mod=1e6, density=1e3, zeta_internal=1.0, zeta_global=0.01, cte=0.0)]
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    m.ijk = np.stack([x.ravel(), y.ravel(), z.ravel() + lift], axis=1).astype(np.int32)
    m.mat = np.zeros(n ** 3, np.int32)
    m.gravity_mult = np.float32(1.0)
    m.floor_enabled = 1
    m.stop_condition_type, m.stop_condition_value, m.equilibrium_mode = sc, value, eq
    return m


@pytest.mark.parametrize("value,stop", [(10.0, 1000), (1500.4, 1501), (1500.6, 1502)])
def test_stop_after_steps(value, stop):
    # steps > (int)(value + 0.5), never before 2 * HISTORY_SIZE = 1000 steps
    o = OSim(_drop_cube(1, value))
    assert o.step(100000) == stop and o.status().state == 2 - 1


@pytest.mark.parametrize("sc,value", [(2, 0.05), (3, 0.2)])
def test_stop_after_time_and_temperature_cycles(sc, value):
    m = _drop_cube(sc, value)
    m.temp_period = 0.25
    limit = value if sc == 2 else m.temp_period * value
    o = OSim(m)
    n = o.step(100000)
    t = o.status().current_time
    assert t > limit and n >= 1000
    before = OSim(m)
    before.step(n - 1)
    assert before.status().current_time <= limit or n == 1000


def test_free_fall_histories_closed_form():
    # one voxel falling without floor or drag: p_n = n*m*g*dt, kinetic energy p^2/(2m),
    # largest move per step |p|/m*dt
    m = _drop_cube(5, 1e-9, n=1)
    m.materials[0].zeta_global = 0.0
    m.floor_enabled = 0
    o = OSim(m)
    o.step(300)
    dt = float(np.float32(o.status().dt))
    mass = 1e3 * 0.01 ** 3
    k, t, mm = o.voxcad_history()
    p = np.abs(o.state()["lin_mom"][0, 2])
    assert p == pytest.approx(300 * mass * 9.80665 * dt, rel=1e-5)
    assert k[0] == pytest.approx(p * p / (2 * mass), rel=1e-5)
    assert k[1] == pytest.approx((299 * mass * 9.80665 * dt) ** 2 / (2 * mass), rel=1e-5)
    assert (t == -1).all() and (mm == -1).all()          # only what the condition needs
    m.stop_condition_type = 6
    o = OSim(m)
    o.step(300)
    k, t, mm = o.voxcad_history()
    rec_dt = o.status().recommended_dt * m.dt_frac      # the max-move history uses DtFrac * OptimalDt in double
    assert mm[0] == pytest.approx(300 * 9.80665 * dt * rec_dt, rel=1e-5) and (k == -1).all()


def test_strain_energy_of_a_stretched_bond():
    # voxel 0 fixed, voxel 1 pulled along +x by F: at rest the tensile strain energy is
    # F^2/(2*a1) with a1 = E*s, the bond's axial stiffness
    m, second, _ = two_voxels("X_POS", ALL & ~TX, force=(0.5, 0, 0))
    m.stop_condition_type, m.stop_condition_value = 4, 1e-12
    m.gravity_mult = np.float32(0.0)
    m.max_steps = 4000
    o = OSim(m)
    o.step(4000)
    k, t, _ = o.voxcad_history()
    assert k[0] < 1e-12
    assert t[0] - k[0] == pytest.approx(0.5 ** 2 / (2 * 1e6 * 0.001), rel=1e-3)


def test_equilibrium_zeroes_motion_when_kinetic_energy_falls():
    # after each step the motion is zeroed (kinetic energy exactly 0) iff the kinetic
    # energy of the three previous steps summed below that of the three before them
    o = OSim(_drop_cube(0, 0.0, eq=1, lift=1))
    ke = []
    for _ in range(3000):
        o.step(1)
        ke.append(o.voxcad_history()[0][0])
    hist = [-1.0] * 6
    zeroed = 0
    for n, e in enumerate(ke):
        k = hist[-1:-7:-1]      # newest first
        dec = sum(k[:3]) < sum(k[3:6]) and 0 not in k
        assert (e == 0.0) == dec, n
        zeroed += dec
        hist.append(e)
    assert zeroed > 10


def test_min_kinetic_energy_stop_brute_force():
    # stop at the first step n >= 1000 at which the kinetic energy per voxel at steps n,
    # n-50, ..., n-450 is at most 0.001*value
    value = 1e-3
    o = OSim(_drop_cube(5, value))
    ke = []
    while o.status().state == 0:
        if o.step(1) == 0:
            break
        ke.append(o.voxcad_history()[0][0])
    nv = 27
    first = next(n for n in range(1000, len(ke) + 1)
                 if all(ke[n - 1 - a] / nv <= 0.001 * value for a in range(0, 500, 50)))
    assert len(ke) == first


# ---- VX3 cilia and force field (VX3_VoxelyzeKernel.cu:846-859, VX3_Voxel.cu:178-184,376)

def _free_voxel(n=1, **kw):
    m = Model(voxel_size=0.01)
    m.materials = [Material(elastic_mod=1e6, density=1e3, zeta_internal=1.0, zeta_global=0.0, **kw)]
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    m.ijk = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    m.mat = np.zeros(n ** 3, np.int32)
    return m


def test_cilia_push_a_free_voxel_after_switch_on():
    # p = n * F * Cilia * dt once t >= TurnOnCiliaAfterThisManySeconds, nothing before
    m = _free_voxel(cilia=2.5, cilia_on_after=0.0)
    m.enable_cilia = 1
    m.base_cilia = np.array([[1e-3, 0.0, 0.0]])
    m.shift_cilia = np.array([[5e-3, 5e-3, 5e-3]])      # scaled by the signal, which stays 0
    o = OSim(m)
    o.step(100)
    dt = float(np.float32(o.status().dt))
    p = o.state()["lin_mom"][0]
    assert p[0] == pytest.approx(100 * 1e-3 * 2.5 * dt, rel=1e-9) and p[1] == 0 and p[2] == 0
    m.materials[0].cilia_on_after = 50.5 * dt
    o = OSim(m)
    o.step(100)
    assert o.state()["lin_mom"][0][0] == pytest.approx(49 * 1e-3 * 2.5 * dt, rel=1e-9)
    m.enable_cilia = 0                                  # EnableCilia off: no force
    o = OSim(m)
    o.step(10)
    assert (o.state()["lin_mom"][0] == 0).all()


def test_cilia_rotate_with_the_voxel():
    # a voxel held at 90 degrees about z pushes along +y for a base force along +x
    m = _free_voxel(cilia=1.0)
    m.enable_cilia = 1
    m.base_cilia = np.array([[1e-3, 0.0, 0.0]])
    m.externals = [Ext(0, dof_fixed=RX | RY | RZ, rotate=(0.0, 0.0, np.pi / 2))]
    o = OSim(m)
    o.step(102)    # the first step still starts from the identity orientation
    dt = float(np.float32(o.status().dt))
    p = o.state()["lin_mom"][0]
    assert p[1] == pytest.approx(101 * 1e-3 * dt, rel=1e-9)
    assert p[0] == pytest.approx(1e-3 * dt, rel=1e-9)


def test_cilia_act_on_surface_voxels_only():
    # the centre voxel of a 3x3x3 cube is interior: its cilia do nothing
    m = _free_voxel(n=3, cilia=1.0)
    m.enable_cilia = 1
    base = np.zeros((27, 3))
    base[13] = (1e-3, 0, 0)
    m.base_cilia = base
    o = OSim(m)
    o.step(200)
    assert np.abs(o.state()["lin_mom"]).max() == 0
    base[13] = 0
    base[0] = (1e-3, 0, 0)
    m.base_cilia = base
    o = OSim(m)
    o.step(200)
    dt = float(np.float32(o.status().dt))
    assert o.state()["lin_mom"][:, 0].sum() == pytest.approx(200 * 1e-3 * dt, rel=1e-3)


def test_force_field_constant_and_time_dependent():
    # x: constant c -> p = n*c*dt; y: the variable t -> p = dt * sum of the step start times
    m = _free_voxel()
    m.force_field = [[(MT["mtCONST"], 2e-3), (MT["mtEND"], 0.0)], [(MT["mtVAR"], 4.0), (MT["mtEND"], 0.0)], []]
    o = OSim(m)
    dt = None
    ts = []
    for _ in range(50):
        ts.append(o.status().current_time)
        o.step(1)
        dt = float(np.float32(o.status().dt))
    p = o.state()["lin_mom"][0]
    assert p[0] == pytest.approx(50 * 2e-3 * dt, rel=1e-9)
    assert p[1] == pytest.approx(dt * sum(ts), rel=1e-9) and p[1] > 0
    assert p[2] == 0


# ---- VX3 detach (gpu_update_detach, VX3_VoxelyzeKernel.cu:946-967)

def _shrinking_core(detach):
    """3x3x3 cube whose centre voxel contracts with the actuation until its six links
    pass their failure strain (1e4 Pa / 1e6 Pa)."""
    m = Model(voxel_size=0.01)
    kw = dict(elastic_mod=1e6, density=1e3, u_static=1.0, u_dynamic=0.5, zeta_internal=1.0, zeta_global=0.01,
              zeta_collision=0.8, mat_model=1, fail_stress=1e4)
    m.materials = [Material(cte=0.0, **kw), Material(cte=0.02, **kw)]
    z, y, x = np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij")
    m.ijk = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    m.mat = np.zeros(27, np.int32)
    m.mat[13] = 1
    m.gravity_mult = np.float32(1.0)
    m.floor_enabled = 1
    m.temp_enabled = m.vary_temp_enabled = 1
    m.temp_amplitude, m.temp_period = 10.0, 0.1
    m.enable_detach = detach
    return m


def test_detach_frees_failed_links_and_surface_follows_a_step_later():
    INTERIOR = 2       # the engine's "SURFACE" bit: set while all six links are present
    o = OSim(_shrinking_core(1))
    fail = 1e4 / 1e6
    k = None
    for n in range(1, 3000):
        o.step(1)
        ms = o.link_state()["max_strain"]
        if k is None and (ms > fail).any():
            k = n
            frozen = o.link_state()["strain"].copy()
            assert o.state()["flags"][13] & INTERIOR      # not yet: the surface follows at the next step
        elif k is not None:
            assert not (o.state()["flags"][13] & INTERIOR)
            failed = frozen.copy()
            gone = o.link_state()["max_strain"] > fail
            assert np.array_equal(o.link_state()["strain"][gone], failed[gone])   # detached links no longer update
            if n > k + 20:
                break
    assert k is not None and gone.sum() == 6                # the centre's six links
    # without detach the failed links stay in place
    o2 = OSim(_shrinking_core(0))
    o2.step(k + 20)
    assert o2.state()["flags"][13] & INTERIOR


# ---- VX3 secondary experiment (VX3_VoxelyzeKernel.cu:336-350, 365-397)

def _layered_cube(remove_after, reinit_after=1e9):
    """3x3x3 cube on the floor whose top layer is a second material, removed once the
    step start time passes remove_after."""
    m = Model(voxel_size=0.01)
    kw = dict(elastic_mod=1e6, density=1e3, u_static=1.0, u_dynamic=0.5, zeta_internal=1.0, zeta_global=0.01,
              zeta_collision=0.8)
    m.materials = [Material(**kw), Material(remove_after=remove_after, **kw)]
    z, y, x = np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij")
    m.ijk = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    m.mat = (z.ravel() == 2).astype(np.int32)
    m.gravity_mult = np.float32(1.0)
    m.floor_enabled = 1
    m.temp_period = 0.002
    m.secondary_experiment = 1
    m.reinit_after = reinit_after
    return m


def test_removed_voxels_freeze_and_lose_their_links():
    REMOVED, INTERIOR = 16, 2
    o = OSim(_layered_cube(1e-9))
    dt = None
    n_removed_at = None
    for n in range(1, 200):
        t_start = o.status().current_time
        o.step(1)
        dt = o.status().dt
        fl = o.state()["flags"]
        top = np.arange(18, 27)
        if n_removed_at is None and (fl[top] & REMOVED).any():
            n_removed_at = n
            assert t_start > 0 and (fl[top] & REMOVED).all() and not (fl[:18] & REMOVED).any()
            frozen = o.state()["pos"][top].copy()
            assert fl[13] & INTERIOR          # the middle voxel loses its top link; the surface follows next step
        elif n_removed_at is not None:
            assert np.array_equal(o.state()["pos"][top], frozen)
            assert not (fl[13] & INTERIOR)
    assert n_removed_at == 2                 # the first step starts at t = 0, not after 1e-9
    o = OSim(_layered_cube(0.0))             # a removal time of 0 means never
    o.step(50)
    assert not (o.state()["flags"] & REMOVED).any()
    # the removal time counts: a later one removes later
    o = OSim(_layered_cube(100.5 * float(np.float32(dt))))
    o.step(100)
    assert not (o.state()["flags"] & REMOVED).any()
    o.step(2)
    assert (o.state()["flags"][18:] & REMOVED).all()


def test_initial_position_reset():
    # once the step start time passes reinit_after, the initial center of mass becomes the
    # last sampled one (sampled every CycleStep = int(TempPeriod/dt) steps) and the total
    # distance counts from the positions at that moment
    probe = OSim(_layered_cube(1e9))
    probe.step(1)
    dt = float(np.float32(probe.status().dt))
    cycle = int(0.002 / dt)
    m = _layered_cube(1e9, reinit_after=(cycle + 10.5) * dt)
    m.fitness_tokens = [(MT["mtVAR"], 2.0), (MT["mtEND"], 0.0)]      # z displacement of the center of mass
    twin = OSim(m)
    twin.step(cycle)
    sampled = np.array(twin.status().current_com[:])
    o = OSim(m)
    o.step(cycle + 11)
    assert np.array_equal(np.array(o.status().initial_com[:]), np.array(OSim(m).status().initial_com[:]))
    o.step(1)                                                         # the reset, at the end of this step
    s2 = o.status()
    assert np.array_equal(np.array(s2.initial_com[:]), sampled)
    pos2 = o.state()["pos"].copy()
    o.step(300)
    s3 = o.status()
    d = np.linalg.norm(o.state()["pos"] - pos2, axis=1).sum()
    assert s3.total_distance_of_all_voxels == pytest.approx(d, rel=1e-12)
    assert s3.fitness == pytest.approx(s3.current_com[2] - sampled[2], rel=1e-12)


# ---- VX3 signals (VX3_Voxel.cu:279-348)

def _chain(n=6, period=10.0, **kw):
    """A weightless row of n voxels along x; voxel 0 is a pacemaker."""
    m = Model(voxel_size=0.01)
    base = dict(elastic_mod=1e6, density=1e3, zeta_internal=1.0, zeta_global=0.0)
    m.materials = [Material(is_pacemaker=1, pacemaker_period=period, **base, **kw), Material(**base, **kw)]
    m.ijk = np.stack([np.arange(n), np.zeros(n), np.zeros(n)], axis=1).astype(np.int32)
    m.mat = np.array([0] + [1] * (n - 1), np.int32)
    m.enable_signals = 1
    return m


def test_signal_travels_down_a_chain():
    # a pulse of 100 starts at the pacemaker; a receiver shows the value at once but
    # passes it on only signalTimeDelay (0.03 s, to within a step) later, 0.9 of it; the
    # refractory period keeps it from echoing back
    o = OSim(_chain())
    first, value = {}, {}
    times = []
    for n in range(1, 4500):
        times.append(o.status().current_time)
        o.step(1)
        sig = o.local_signal()
        for k in range(6):
            if k not in first and sig[k] > 0:
                first[k], value[k] = n, sig[k]
    dt = float(np.float32(o.status().dt))
    assert value[0] == pytest.approx(90.0)                   # 100, decayed once in the same step
    for k in range(1, 6):
        assert value[k] == pytest.approx(100 * 0.9 ** k, rel=1e-12)
    assert first[1] == 2                                       # the pacemaker posts at step 2
    for k in range(2, 6):
        hop = (first[k] - first[k - 1]) * dt
        assert 0.03 <= hop + 1e-12 and hop <= 0.03 + 2 * dt + 1e-12


def test_pacemaker_period_and_decay():
    # a lone pacemaker fires every period; in between its local signal decays by 0.9 every 0.01 s
    m = _chain(n=1, period=0.05)
    o = OSim(m)
    sig, t0 = [], []
    for _ in range(2500):
        t0.append(o.status().current_time)
        o.step(1)
        sig.append(o.local_signal()[0])
    sig = np.array(sig)
    fires = [i for i in range(1, len(sig)) if sig[i] > sig[i - 1]]
    gaps = np.diff([t0[i] for i in [0] + fires])
    assert len(fires) >= 2 and np.all(gaps >= 0.05 - 1e-12) and np.all(gaps <= 0.05 + 2e-4)
    assert sig.min() > 0 and sig.max() == 100.0
    assert np.all(np.isin(np.round(sig[1:fires[0]], 9), np.round(90 * 0.9 ** np.arange(8), 9)))   # 0.9 per 0.01 s


def test_signal_scales_the_cilia_shift_force():
    # the cilia force is RotateVec3D(base + localSignal*shift)*Cilia with the signal as
    # of the step's start
    m = _chain(n=1, period=0.05, cilia=2.0)
    m.enable_cilia = 1
    m.base_cilia = np.array([[0.0, 0.0, 0.0]])
    m.shift_cilia = np.array([[1e-5, 0.0, 0.0]])
    o = OSim(m)
    used = 0.0
    for _ in range(400):
        used += o.local_signal()[0]
        o.step(1)
    dt = float(np.float32(o.status().dt))
    assert o.state()["lin_mom"][0][0] == pytest.approx(used * 1e-5 * 2.0 * dt, rel=1e-9) and used > 0


# ---- VX3 target contacts (handle_collision_attachment, VX3_VoxelyzeKernel.cu:682-722)

def _target_drop():
    from gpuvoxels_amd import workloads
    m = workloads.drop(3, gap=1, shift=1, steps=0)
    m.materials = [workloads.base_material(cte=0.0), workloads.base_material(cte=0.0, is_target=1)]
    m.mat = (m.ijk[:, 2] >= 3).astype(np.int32)      # the upper cube is the target
    return m


def test_target_hits_brute_force():
    # each step counts the watched pairs of one target and one other voxel whose
    # centres are within 0.625 x (size_a + size_b) at the step's start
    m = _target_drop()
    target = m.mat == 1
    o = OSim(m)
    o.step(2000)
    checked = 0
    for _ in range(500):
        pos, pairs, h0 = o.state()["pos"].copy(), o.collision_pairs().copy(), o.status().collision_count
        o.step(1)
        if not np.array_equal(pairs, o.collision_pairs()):
            continue          # the watch list was rebuilt during the step
        want = sum(1 for a, b in pairs if target[a] != target[b]
                   and np.linalg.norm(pos[a] - pos[b]) <= (0.01 + 0.01) * 0.625)
        assert o.status().collision_count - h0 == want
        checked += want > 0
    assert checked > 50


def test_target_hit_sends_a_signal_and_feeds_the_fitness():
    m = _target_drop()
    m.enable_signals = 1
    m.fitness_tokens = [(MT["mtVAR"], 3.0), (MT["mtEND"], 0.0)]     # hit
    o = OSim(m)
    o.step(2600)
    st = o.status()
    assert st.collision_count > 0 and st.fitness == st.collision_count
    sig = o.local_signal()
    assert sig[m.mat == 0].max() > 0 and sig[m.mat == 1].max() == 0   # only non-target voxels receive


# ---- VX3 attach (handle_collision_attachment, VX3_VoxelyzeKernel.cu:724-826)

def _sticky_bars(n=7, sticky=(1, 1), gap=2, guard=500, conditions=None, dt=None):
    """Two n-voxel bars along x, one on the floor and one `gap` voxels above it, falling
    onto it with collisions off; material 0 for the lower bar, 1 (or 0) for the upper."""
    m = Model(voxel_size=0.01)
    kw = dict(elastic_mod=1e6, density=1e3, u_static=1.0, u_dynamic=0.5, zeta_internal=1.0, zeta_global=0.01,
              zeta_collision=0.8)
    m.materials = [Material(sticky=sticky[0], **kw), Material(sticky=sticky[1], **kw)]
    x = np.arange(n)
    m.ijk = np.concatenate([np.stack([x, 0 * x, 0 * x], 1), np.stack([x, 0 * x, 0 * x + gap], 1)]).astype(np.int32)
    m.mat = np.array([0] * n + [1 if sticky[0] != sticky[1] else 0] * n, np.int32)
    m.gravity_mult = np.float32(1.0)
    m.floor_enabled = 1
    m.collisions_enabled = 0
    m.enable_attach = 1
    m.safety_guard = guard
    if conditions:
        m.attach_conditions = conditions
    if dt:
        m.dt_fixed = dt
    return m


def _links_between(o, n):
    vneg, vpos, axis = o.topology()[:3]
    return sorted((int(a), int(b), int(ax)) for a, b, ax in zip(vneg, vpos, axis) if (a < n) != (b < n))


def test_attach_bars_take_pairs_in_order_and_skip_those_within_five_links():
    # the upper bar falls until the vertical pairs come within 0.625 x (1 + 1) voxels; taken in
    # order, (B0, A0) links first; (B1, A1) and (B2, A2) are then within 3 and 5 links, (B3, A3)
    # 7 links apart attaches, and so on: new links at 0, 3 and 6, each along z from the lower
    # voxel (negative end) to the upper one
    n = 7
    o = OSim(_sticky_bars(n))
    nl0 = o.nl
    o.step(3000)
    assert o.nl == nl0 + 3
    assert _links_between(o, n) == [(0, n, 2), (3, n + 3, 2), (6, n + 6, 2)]


def test_attach_needs_one_sticky_material_and_its_conditions():
    for m in (_sticky_bars(sticky=(0, 0)),                      # not sticky
              _sticky_bars(sticky=(1, 2)),                      # two sticky materials
              _sticky_bars(conditions=[[(MT["mtCONST"], 1.0), (MT["mtEND"], 0.0)], [],
                                       [(MT["mtCONST"], -1.0), (MT["mtEND"], 0.0)], [], []])):   # one condition fails
        o = OSim(m)
        nl0 = o.nl
        o.step(3000)
        assert o.nl == nl0
    # a condition of the voxel height, 11.5 mm - z > 0: the upper bar's voxels qualify only
    # once they have fallen below it, later than the 12.5 mm envelope alone; the links are the same
    cond = [[(MT["mtVAR"], 2.0), (MT["mtCONST"], 0.0115), (MT["mtSUB"], 0.0), (MT["mtEND"], 0.0)], [], [], [], []]
    made = []
    for c in (None, cond):
        o = OSim(_sticky_bars(conditions=c))
        k, nl0 = 0, o.nl
        while o.nl == nl0 and k < 3000:
            o.step(1)
            k += 1
        made.append((k, min(o.state()["pos"][7:, 2])))
        o.step(3000 - k)
        assert _links_between(o, 7) == [(0, 7, 2), (3, 10, 2), (6, 13, 2)]
    assert made[1][0] > made[0][0] and made[1][1] < 0.0115 < made[0][1]


def test_new_link_carries_no_force_then_one_hundredth():
    # three runs in lockstep: no guard, the guard, and no sticky material.  They are equal
    # through the step that makes the link (it carries no force there); at the next step
    # the upper voxel's momentum change is G + L without the guard and G + L/100 with it,
    # G being the change without the link
    sims = [OSim(_sticky_bars(n=1, guard=0, dt=1e-4)), OSim(_sticky_bars(n=1, guard=500, dt=1e-4)),
            OSim(_sticky_bars(n=1, sticky=(0, 0), dt=1e-4))]
    made = None
    for k in range(4000):
        prev = [o.state() for o in sims]
        for o in sims:
            o.step(1)
        if sims[1].nl == 1:
            made = k
            break
    assert made is not None and sims[0].nl == 1 and sims[2].nl == 0
    now = [o.state() for o in sims]
    assert np.array_equal(now[0]["pos"], now[2]["pos"]) and np.array_equal(now[1]["lin_mom"], now[2]["lin_mom"])
    for o in sims:
        o.step(1)
    after = [o.state() for o in sims]
    X, Y, G = [a["lin_mom"][1, 2] - b["lin_mom"][1, 2] for a, b in zip(after, now)]
    assert abs(X - G) > 1e3 * abs(G) * 1e-12
    assert (Y - G) == pytest.approx(0.01 * (X - G), rel=1e-9)


def _close_targets():
    """Three material-fixed target voxels on the x axis at 0, 2 and 5 voxel sizes and a
    free non-target voxel far away; MaxDistInVoxelLengthsToCountAsPair = 3, center of
    mass sampled every 10 steps (TempPeriod / dt = 10)."""
    s = 0.01
    m = Model(voxel_size=s)
    m.materials = [Material(elastic_mod=1e6, density=1000.0, fixed=1, is_target=1, matid=1),
                   Material(elastic_mod=1e6, density=1000.0, matid=2)]
    m.ijk = np.array([[0, 0, 0], [2, 0, 0], [5, 0, 0], [9, 9, 9]], np.int32)
    m.mat = np.array([0, 0, 0, 1], np.int32)
    m.dt_fixed = 1e-4
    m.temp_period = 1e-3
    m.max_dist_close_pairs = 3.0
    return m


def test_target_closeness_and_close_pairs_closed_form():
    """computeTargetCloseness (VX3_VoxelyzeKernel.cu:545-562): at the sample, pairs of
    target voxels closer than 3 voxel sizes are counted (only 0-2: the 3-size pair 2-5
    is not closer than 3) and 1/distance is summed over all three pairs; the stop
    program numClosePairs > 0 stops the run after the sample step."""
    m = _close_targets()
    s = m.voxel_size
    o = OSim(m)
    o.step(9)
    assert o.status().num_close_pairs == 0 and o.status().target_closeness == 0
    o.step(1)
    st = o.status()
    assert st.num_close_pairs == 1
    assert st.target_closeness == pytest.approx(1 / (2 * s) + 1 / (5 * s) + 1 / (3 * s), rel=1e-12)
    m.stop_tokens = [(MT["mtCONST"], 0.0), (MT["mtVAR"], 7.0), (MT["mtSUB"], 0.0), (MT["mtEND"], 0.0)]
    o2 = OSim(m)
    assert o2.step(100) == 10 and o2.status().state == 1


# ---- VX3 topology changes with collisions: the nearby lists follow the current links

def _brute_pairs(pos, flags, voxel_links, voxel_size, hops=5, br=0.75, wd=1.0):
    """CVoxelyze::regenerateCollisions (Voxelyze.cpp:726-751) by brute force: surface
    voxel pairs i < j within the watch radius, unless j is within `hops` links of i over
    the links present (generateNearby on the current topology)."""
    INTERIOR = 2
    n = len(pos)
    adj = [[] for _ in range(n)]
    for v in range(n):
        for d in range(6):
            l = voxel_links[v, d]
            if l >= 0:
                adj[v].append(l)
    # voxel_links holds link indices; map each to its two voxels
    ends = {}
    for v in range(n):
        for d in range(6):
            l = voxel_links[v, d]
            if l >= 0:
                ends.setdefault(l, []).append(v)
    nb = [[u for l in adj[v] for u in ends[l] if u != v] for v in range(n)]
    wr = np.float32(voxel_size * np.float32(2 * np.float32(br) + np.float32(wd)))
    thr = float(np.float32(wr * wr))
    out = []
    for i in range(n):
        if flags[i] & INTERIOR:
            continue
        near, frontier = {i}, [i]
        for _ in range(hops):
            frontier = [u for v in frontier for u in nb[v] if u not in near and not near.add(u)]
        for j in range(i + 1, n):
            if flags[j] & INTERIOR or j in near:
                continue
            if float(np.sum((pos[i] - pos[j]) ** 2)) <= thr:
                out.append((i, j))
    return np.array(out, np.int32).reshape(-1, 2)


def test_detach_with_collisions_rebuilds_nearby_from_current_links():
    """The shrinking core with collisions: once the centre voxel's links fail it is a
    free surface voxel, no longer within five links of the shell, so at the next step
    it is watched against its 26 neighbours; the pair list equals the brute-force
    rule on the positions of the regeneration and the links present then."""
    m = _shrinking_core(1)
    m.collisions_enabled = 1
    o = OSim(m)
    fail = 1e4 / 1e6
    for n in range(1, 3000):
        o.step(1)
        if (o.link_state()["max_strain"] > fail).any():
            break
    before = o.collision_pairs()
    assert not any(13 in p for p in before)              # still interior, still linked
    st = o.state()                                       # positions the next step regenerates from
    o.step(1)
    vl = o.topology()[3]
    got = o.collision_pairs()
    want = _brute_pairs(st["pos"], o.state()["flags"], vl, m.voxel_size)
    assert np.array_equal(got, want)
    assert sum(13 in p for p in got) == 26


def test_kinetic_friction_is_discontinuous_at_rest():
    """CVX_Voxel::floorForce (old/VX_Voxel.cpp:272-297): off the static latch, a voxel in
    floor contact is slowed by mu_k * N along its normalized horizontal velocity, so the
    force does not shrink with the velocity.  A voxel at rest gets none; the same voxel
    started at 1e-15 m/s gets the full kinetic force and within a few steps moves at
    more than 1e6 times that speed.  Rounding differences of that size are why FP32
    state (or FMA-contracted FP64) cannot follow these trajectories closely."""
    m = Model(voxel_size=0.01)
    m.materials = [Material(elastic_mod=1e6, density=1000.0, u_static=1.0, u_dynamic=0.5, zeta_internal=1.0,
                            zeta_global=0.01, zeta_collision=0.8)]
    m.ijk = np.array([[0, 0, 0]], np.int32)
    m.mat = np.zeros(1, np.int32)
    m.gravity_mult = float(np.float32(1.0))
    m.floor_enabled = 1
    a = OSim(m)
    a.step(10)
    assert a.state()["lin_mom"][0, 0] == 0.0
    m.init_velocity = np.array([[1e-15, 0.0, 0.0]])
    b = OSim(m)
    b.step(10)
    mass = 1000.0 * 0.01 ** 3
    assert abs(b.state()["lin_mom"][0, 0]) / mass > 1e6 * 1e-15
