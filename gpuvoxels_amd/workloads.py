"""Synthetic, seeded workloads (the benchmark and parity configurations).

* ``cube``      6x6x6 solid cube, one actuated material (single-sim config).
* ``robots``    uneven robots on a 12^3 grid: Bernoulli(0.5) occupancy, largest
                face-connected component, one of four materials per voxel
                (soft passive, stiff passive, actuated phase 0, actuated phase
                half a cycle).  The component rule restates
                Experiments/framework/voxelyze/helper.py:1-17 (scipy ``label`` with
                its default face connectivity; the first label with the strictly
                largest count wins) and the body layout of
                Experiments/framework/voxelyze/workflow.py:45-60 (array [z, y, x],
                x fastest in the structure order).
* ``slab``      64x64x32 bilinear slab with a four-phase checkerboard.
* ``arena_balls`` 2048 radius-4 balls thrown into a 400x400 arena (collisions).
"""
import numpy as np

from . import _abi
from .model import Material, Model, FITNESS_XY_DISTANCE, FITNESS_DISTANCE_3D


def gravity_mult(grav_acc: float = -9.81) -> float:
    """CVX_Sim::SetGravityAccel: Vx.setGravity(-grav/9.80665) with a float argument
    (Voxelyze3/src/VXA/VX_Sim.cpp:485)."""
    g = np.float32(grav_acc)
    return float(np.float32(-float(g) / 9.80665))


def _physics(m: Model, steps: int):
    m.gravity_mult = gravity_mult(-9.81)
    m.gravity_acc = -9.81
    m.floor_enabled = 1
    m.temp_enabled = 1
    m.vary_temp_enabled = 1
    m.temp_base = 25.0
    m.temp_amplitude = 10.0       # CTE * A = 0.1
    m.temp_period = 0.5           # 2 Hz
    m.enable_expansion = 0
    m.dt_frac = 0.9
    m.volume_effects = 1
    m.collisions_enabled = 0
    m.max_steps = steps
    m.fitness_tokens = list(FITNESS_XY_DISTANCE)


def base_material(**kw) -> Material:
    d = dict(elastic_mod=1e6, density=1000.0, poissons_ratio=0.35, cte=0.01, u_static=1.0, u_dynamic=0.5,
             zeta_internal=1.0, zeta_collision=0.8, zeta_global=0.01, matid=1)
    d.update(kw)
    return Material(**d)


def cube(n: int = 6, steps: int = 10000, voxel_size: float = 0.01) -> Model:
    """Solid n^3 cube of one actuated material (E=1e6, nu=0.35, rho=1000, CTE 0.01)."""
    m = Model(voxel_size=voxel_size, name="cube%d" % n)
    m.materials = [base_material()]
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    m.ijk = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    m.mat = np.zeros(n ** 3, np.int32)
    m.phase = np.zeros(n ** 3)
    _physics(m, steps)
    m.fitness_tokens = list(FITNESS_DISTANCE_3D)   # a symmetric cube does not walk: score its settling
    return m


def largest_component(body: np.ndarray) -> np.ndarray:
    """Keeps the largest face-connected component of a boolean [z, y, x] body."""
    from scipy.ndimage import label
    labeled, total = label(body)
    best_id, best_count = 0, 0
    counts = np.bincount(labeled.ravel(), minlength=total + 1)
    for i in range(1, total + 1):
        if counts[i] > best_count:
            best_count, best_id = counts[i], i
    return labeled == best_id if best_id > 0 else np.zeros_like(body, dtype=bool)


ROBOT_MATERIALS = [
    dict(elastic_mod=1e5, cte=0.0, matid=1),   # soft passive
    dict(elastic_mod=1e7, cte=0.0, matid=2),   # stiff passive
    dict(elastic_mod=1e6, cte=0.01, matid=3),  # actuated, phase 0
    dict(elastic_mod=1e6, cte=0.01, matid=4),  # actuated, phase half a cycle
]
ROBOT_PHASE = [0.0, 0.0, 0.0, 0.5]


def robot(rng: np.random.Generator, grid: int = 12, p: float = 0.5, steps: int = 20000, voxel_size: float = 0.01,
          name: str = "robot") -> Model:
    occ = rng.random((grid, grid, grid)) < p
    body = largest_component(occ)
    mats = rng.integers(0, 4, size=body.shape)
    zyx = np.argwhere(body)                       # lexicographic (z, y, x): x fastest
    m = Model(voxel_size=voxel_size, name=name)
    m.materials = [base_material(**kw) for kw in ROBOT_MATERIALS]
    m.ijk = zyx[:, ::-1].astype(np.int32).copy()
    m.mat = mats[body].astype(np.int32)
    m.phase = np.array(ROBOT_PHASE)[m.mat]
    _physics(m, steps)
    return m


def robots(n: int = 512, seed: int = 0, grid: int = 12, steps: int = 20000, voxel_size: float = 0.01):
    rng = np.random.default_rng(seed)
    return [robot(rng, grid=grid, steps=steps, voxel_size=voxel_size, name="robot_%04d" % i) for i in range(n)]


def slab(nx: int = 64, ny: int = 64, nz: int = 32, steps: int = 5000, voxel_size: float = 0.01,
         collisions: bool = True) -> Model:
    """Bilinear slab (E=1e6 to yield strain 0.05, then 1e5) with a 4-phase checkerboard."""
    m = Model(voxel_size=voxel_size, name="slab")
    m.materials = [base_material(mat_model=_abi.MDL_BILINEAR, plastic_mod=1e5, yield_stress=5e4, fail_stress=1e7,
                                 matid=k + 1) for k in range(4)]
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    m.ijk = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    m.mat = (((x // 4) % 2) * 2 + ((y // 4) % 2)).ravel().astype(np.int32)
    m.phase = np.array([0.0, 0.25, 0.5, 0.75])[m.mat]
    _physics(m, steps)
    m.collisions_enabled = 1 if collisions else 0
    return m


def drop(n: int = 3, gap: int = 1, shift: int = 1, steps: int = 2000, voxel_size: float = 0.01) -> Model:
    """Two detached n^3 cubes in one simulation: the upper one, `gap` voxels above and
    `shift` voxels across, falls onto the lower one (voxel-voxel contacts)."""
    m = Model(voxel_size=voxel_size, name="drop")
    m.materials = [base_material(cte=0.0)]
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    lower = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    upper = lower + np.array([shift, shift, n + gap])
    cells = np.concatenate([lower, upper])
    order = np.lexsort((cells[:, 0], cells[:, 1], cells[:, 2]))    # structure-index order
    m.ijk = cells[order].astype(np.int32)
    m.mat = np.zeros(len(cells), np.int32)
    _physics(m, steps)
    m.vary_temp_enabled = 0
    m.collisions_enabled = 1
    return m


def balls(n_side: int = 4, layers: int = 2, radius: int = 4, gap: float = 1.5, steps: int = 3000,
          voxel_size: float = 0.01, seed: int = 0) -> Model:
    """Many voxel balls in one simulation (the many-body collision configuration):
    n_side x n_side balls of the given radius per layer, centres `2*radius + 1 + gap`
    voxels apart, each layer shifted by half a spacing and raised by one spacing, so
    upper balls fall into the gaps of the lower layer (ball-ball and floor contacts).
    Ball voxels: lattice cells within `radius` of the centre (257 for radius 4)."""
    m = Model(voxel_size=voxel_size, name="balls")
    m.materials = [base_material(cte=0.0, zeta_global=1e-4)]   # light air drag: the balls fall freely
    r = radius
    d = np.arange(-r, r + 1)
    z, y, x = np.meshgrid(d, d, d, indexing="ij")
    inside = x * x + y * y + z * z <= r * r
    ball = np.stack([x[inside], y[inside], z[inside]], axis=1)
    spacing = 2 * r + 1 + gap
    cells = []
    for layer in range(layers):
        off = 0.5 * spacing * (layer % 2)
        for i in range(n_side):
            for j in range(n_side):
                c = np.array([r + off + i * spacing, r + off + j * spacing, r + layer * spacing])
                cells.append(ball + np.round(c).astype(np.int64))
    cells = np.concatenate(cells)
    order = np.lexsort((cells[:, 0], cells[:, 1], cells[:, 2]))    # structure-index order
    m.ijk = cells[order].astype(np.int32)
    assert len(np.unique(m.ijk, axis=0)) == len(m.ijk), "balls overlap"
    m.mat = np.zeros(len(cells), np.int32)
    _physics(m, steps)
    m.vary_temp_enabled = 0
    m.collisions_enabled = 1
    return m


def ball_offsets(radius: int = 4) -> np.ndarray:
    """Lattice offsets within `radius` of a ball's centre cell (257 for radius 4)."""
    d = np.arange(-radius, radius + 1)
    z, y, x = np.meshgrid(d, d, d, indexing="ij")
    inside = x * x + y * y + z * z <= radius * radius
    return np.stack([x[inside], y[inside], z[inside]], axis=1)


def arena_balls(n_balls: int = 2048, radius: int = 4, arena: int = 400, gap: float = 1.5, sigma_v: float = 0.05,
                steps: int = 3000, voxel_size: float = 0.01, seed: int = 0, z_layers: int = 3) -> Model:
    """The collision-heavy multi-body configuration: `n_balls` detached balls of the
    given radius in one simulation, centres on lattice cells drawn at random (seeded)
    inside an arena x arena footprint and z_layers ball diameters of height, accepted
    when at least 2*radius + 1 + gap voxels from every ball placed before (so balls
    keep >= gap voxels between their surfaces); every voxel of a ball starts with that
    ball's velocity, drawn per ball from N(0, sigma_v) m/s per axis.  Gravity and floor
    on, no actuation; the broadphase period is a batch setting (8 steps for this
    configuration).  Voxels are listed in structure-index order (x fastest)."""
    rng = np.random.default_rng(seed)
    r = radius
    dmin = 2 * r + 1 + gap
    lo, hi = r, arena - 1 - r
    zlo, zhi = r, r + z_layers * (2 * r + 1)
    cell = int(np.ceil(dmin))
    grid = {}
    centres = []
    tries = 0
    while len(centres) < n_balls:
        tries += 1
        if tries > 200 * n_balls:
            raise RuntimeError("arena too small for %d balls" % n_balls)
        c = (int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1)), int(rng.integers(zlo, zhi + 1)))
        k = (c[0] // cell, c[1] // cell, c[2] // cell)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if (q[0] - c[0]) ** 2 + (q[1] - c[1]) ** 2 + (q[2] - c[2]) ** 2 < dmin * dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            grid.setdefault(k, []).append(c)
            centres.append(c)
    vel = rng.normal(0.0, sigma_v, size=(n_balls, 3))
    off = ball_offsets(r)
    cells = (np.array(centres)[:, None, :] + off[None, :, :]).reshape(-1, 3)
    v = np.repeat(vel, len(off), axis=0)
    order = np.lexsort((cells[:, 0], cells[:, 1], cells[:, 2]))    # structure-index order
    m = Model(voxel_size=voxel_size, name="arena_balls")
    m.materials = [base_material(cte=0.0)]
    m.ijk = cells[order].astype(np.int32)
    m.init_velocity = v[order].copy()
    m.mat = np.zeros(len(cells), np.int32)
    _physics(m, steps)
    m.vary_temp_enabled = 0
    m.collisions_enabled = 1
    return m


def serving_request(rng: np.random.Generator, v_min: int = 100, v_max: int = 50000, h_min: int = 2000,
                    h_max: int = 20000, name: str = "req"):
    """One request of the serving stream: a robot built as in the 512-robot workload on a
    grid scaled so that its expected voxel count (47.5% of the cells survive the largest
    component rule) is log-uniform in [v_min, v_max], and a horizon uniform in
    [h_min, h_max] steps.  Returns (model, horizon)."""
    target = float(np.exp(rng.uniform(np.log(v_min), np.log(v_max))))
    grid = max(5, int(round((target / 0.475) ** (1.0 / 3.0))))
    m = robot(rng, grid=grid, steps=0, name=name)
    return m, int(rng.integers(h_min, h_max + 1))


def serving_stream(seconds: float, rate: float, seed: int = 4, **kw):
    """Poisson arrivals at `rate` requests/s over `seconds`: [(arrival time, model, horizon)]."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return out
        m, h = serving_request(rng, name="req%04d" % len(out), **kw)
        out.append((t, m, h))
