"""Synthetic scenes: seeded jittered Kuhn beams and their material fields.

Meshes are generated in memory (they are not routed through a .mesh loader,
which would rescale coordinates by 1/max, reference
src/simulation_state.cpp:64-67). Randomness comes from a counter-based
splitmix64 stream so that every draw depends only on (seed, stream, index)
and is reproducible independent of numpy's generator versions.

Conventions shared with the reference:
  * vertices are rows of V (nV x dim), elements rows of T (nT x dim+1);
  * "hangends" pins the vertices near the MAX end of `axis`
    (reference src/boundary_conditions/fixed_boundary_conditions.h:151-166),
    so the clamped face of every beam is x = x_max;
  * Young's modulus / Poisson ratio map to Lame parameters as in
    reference src/config.h:13-16.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)

# material ids (reference src/config.h:61-68 order)
MATERIALS = {
    "Stable-Neohookean": 0,
    "Neohookean": 1,
    "Corot": 2,
    "FixedCorot": 3,
    "ARAP": 4,
    "Fung": 5,
}

# boundary condition ids (reference src/config.h:70-87 order)
BCS = {"null": 0, "scale": 1, "hang": 2, "hangends": 3, "stretch": 4,
       "bend": 7, "twist": 8, "onepoint": 13, "randomize": 14}


def splitmix64_uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from splitmix64 applied to a per-(seed, stream) counter."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF)
        x = base + (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def enu_to_lame(E, nu):
    """Reference src/config.h:13-16."""
    E = np.asarray(E, dtype=np.float64)
    mu = E / (2.0 * (1.0 + nu))
    la = (E * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return mu, la


_KUHN_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def kuhn_beam(nx: int, ny: int, nz: int, h: float = 0.1, jitter: float = 0.0,
              seed: int = 0):
    """Box of nx*ny*nz cubes of edge h, each split into 6 tets along its main
    diagonal (Kuhn). Vertex (i,j,k) has index (i*(ny+1)+j)*(nz+1)+k. Every
    coordinate is jittered by U(-jitter, jitter)*h except the x coordinate of
    the x = x_max face, which stays exactly on the clamp plane. Tets are
    positively oriented.
    Returns V (nV x 3, float64) and T (nT x 4, int32)."""
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                          indexing="ij")
    V = np.stack([I, J, K], axis=-1).reshape(-1, 3).astype(np.float64) * h
    nV = V.shape[0]
    if jitter > 0:
        u = splitmix64_uniform(seed, 1, nV * 3).reshape(nV, 3)
        d = (2.0 * u - 1.0) * jitter * h
        d[I.reshape(-1) == nx, 0] = 0.0
        V = V + d

    def vid(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.reshape(-1), cj.reshape(-1), ck.reshape(-1)
    tets = []
    for perm in _KUHN_PERMS:
        off = np.zeros((4, 3), dtype=np.int64)
        for s in range(1, 4):
            off[s] = off[s - 1]
            off[s, perm[s - 1]] += 1
        t = np.stack([vid(ci + off[s, 0], cj + off[s, 1], ck + off[s, 2]) for s in range(4)],
                     axis=1)
        sign = np.linalg.det(np.stack([off[1], off[2], off[3]], axis=1).astype(np.float64))
        if sign < 0:
            t = t[:, [0, 2, 1, 3]]
        tets.append(t)
    T = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    return V, T


def tri_rectangle(nx: int, ny: int, h: float = 0.1):
    """2D rectangle of nx*ny square cells, each split into 2 CCW triangles
    along the (i,j)-(i+1,j+1) diagonal. Vertex (i,j) has index i*(ny+1)+j."""
    I, J = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    V = np.stack([I, J], axis=-1).reshape(-1, 2).astype(np.float64) * h

    def vid(i, j):
        return i * (ny + 1) + j

    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ci, cj = ci.reshape(-1), cj.reshape(-1)
    t0 = np.stack([vid(ci, cj), vid(ci + 1, cj), vid(ci + 1, cj + 1)], axis=1)
    t1 = np.stack([vid(ci, cj), vid(ci + 1, cj + 1), vid(ci, cj + 1)], axis=1)
    T = np.stack([t0, t1], axis=1).reshape(-1, 3).astype(np.int32)
    return V, T


def bc_fixed(kind: str, V: np.ndarray, axis: int = 0, ratio: float = 0.01) -> np.ndarray:
    """Pinned-vertex mask of the reference's boundary conditions
    (boundary_condition.cpp:5-21, fixed_boundary_conditions.h:75-166)."""
    t = BCS[kind]
    mn, mx = V.min(axis=0), V.max(axis=0)
    rng = mx[axis] - mn[axis]
    x = V[:, axis]
    g0 = np.nonzero(x < mn[axis] + rng * ratio)[0]
    g1 = np.nonzero((x >= mn[axis] + rng * ratio) & (x > mx[axis] - rng * ratio))[0]
    fixed = np.zeros(V.shape[0], dtype=np.int32)
    if t == 2:
        if g0.size:
            fixed[g0[-1]] = 1
        if g1.size:
            fixed[g1[-1]] = 1
    elif t == 3:
        fixed[g1] = 1
    elif t in (4, 7, 8):
        fixed[g0] = 1
        fixed[g1] = 1
    elif t == 13:
        fixed[0] = 1
    return fixed


@dataclass
class Scene:
    """Everything one simulation needs, in the reference's units and layout."""
    V: np.ndarray
    T: np.ndarray
    mu: np.ndarray
    la: np.ndarray
    mat: np.ndarray
    fixed: np.ndarray
    density: float = 1000.0
    dt: float = 1.0 / 30.0
    gravity: tuple = (0.0, -9.8, 0.0)
    bdf_order: int = 1
    mixed: bool = True
    solver: str = "eigen-pcg-diag"
    itr_tol: float = 1e-4
    max_iters: int = 500
    max_newton: int = 5
    newton_tol: float = 1e-10
    ls_iters: int = 20
    v0: np.ndarray | None = None
    # start positions (nV x dim) when they differ from the rest shape V: the
    # reference's initial_state x_path
    x0: np.ndarray | None = None
    # boundary condition script {"type", "axis", "ratio", "velocity"}; when set,
    # `fixed` holds the pins it implies and the end groups move after every step
    bc: dict | None = None
    name: str = ""
    extra: dict = field(default_factory=dict)

    @property
    def dim(self):
        return self.V.shape[1]

    @property
    def nV(self):
        return self.V.shape[0]

    @property
    def nT(self):
        return self.T.shape[0]


def beam_scene(nx, ny, nz, h=0.1, jitter=0.1, seed=0, E=1e6, nu=0.45,
               E_range=None, material="Stable-Neohookean", mixed=True, **kw) -> Scene:
    """A clamped (x = max face, "hangends" with ratio 0.5h/L, reference
    fixed_boundary_conditions.h:151-166) jittered Kuhn beam. E_range=(a, b)
    draws a per-tet E log-uniform in [a, b] (stream 2 of the seed)."""
    V, T = kuhn_beam(nx, ny, nz, h, jitter, seed)
    nT = T.shape[0]
    if E_range is not None:
        u = splitmix64_uniform(seed, 2, nT)
        Ev = np.exp(math.log(E_range[0]) + u * (math.log(E_range[1]) - math.log(E_range[0])))
    else:
        Ev = np.full(nT, float(E))
    mu, la = enu_to_lame(Ev, nu)
    mat = np.full(nT, MATERIALS[material], dtype=np.int32)
    L = nx * h
    fixed = bc_fixed("hangends", V, axis=0, ratio=0.5 * h / L * 0.8)
    return Scene(V=V, T=T, mu=mu, la=la, mat=mat, fixed=fixed, mixed=mixed,
                 name=f"beam{nx}x{ny}x{nz}", **kw)


def rect2d_scene(nx=40, ny=10, h=0.1, E=1e5, nu=0.45, **kw) -> Scene:
    """2D plain-method rectangle (800 tris for 40x10); the edge at x = x_max is
    pinned with "hangends" (the reference's hangends pins the max end)."""
    V, T = tri_rectangle(nx, ny, h)
    nT = T.shape[0]
    mu, la = enu_to_lame(np.full(nT, float(E)), nu)
    mat = np.zeros(nT, dtype=np.int32)
    fixed = bc_fixed("hangends", V, axis=0, ratio=0.5 * h / (nx * h) * 0.8)
    kw.setdefault("gravity", (0.0, -9.8, 0.0))
    return Scene(V=V, T=T, mu=mu, la=la, mat=mat, fixed=fixed, mixed=False,
                 name=f"rect{nx}x{ny}", **kw)


def scene_from_json(path: str, dim: int = 3) -> Scene:
    """Arrays of a one-object scene JSON, loaded with the reference's rules
    (reference src/simulation_state.cpp:46-366): 3D Medit meshes scaled by
    1/max coefficient, 2D OBJ triangle meshes with unreferenced vertices
    dropped and z truncated; offset, then the DIM x DIM "transformation"
    (column-major) about the centroid; "material_ids" (3D; the 2D branch
    keeps materials[0]) or "material_index"; "initial_velocity";
    "boundary_condition"; "initial_state" x_path / v_path (DMAT start
    positions and velocities, the rest shape staying the file's mesh);
    solver parameters. Multi-object scenes are rejected."""
    import json
    import os

    from . import meshio
    d = json.load(open(path))
    base = os.path.dirname(os.path.abspath(path))
    def res(q):
        # against the scene file's folder, else as given (the reference opens
        # paths relative to the working directory)
        if os.path.isabs(q):
            return q
        here = os.path.join(base, q)
        return here if os.path.exists(here) or not os.path.exists(q) else q
    objs = d.get("objects", [])
    if len(objs) != 1:
        raise ValueError("scene_from_json loads scenes with exactly one object")
    obj = objs[0]
    if dim == 3:
        V, T = meshio.read_mesh(res(obj["path"]))
        V = V / V.max()
    else:
        V, T = meshio.read_obj(res(obj["path"]))
        V, T = meshio.remove_unreferenced(V, T)
        V = V[:, :2].copy()
    off = np.asarray(obj.get("offset", [0, 0, 0]), dtype=np.float64)
    V = V + off[None, :dim]
    if "transformation" in obj:
        A = np.asarray(obj["transformation"], dtype=np.float64).reshape(dim, dim, order="F")
        c = V.sum(axis=0) / V.shape[0]
        V = (V - c) @ A.T + c
    nT = T.shape[0]
    mats = d.get("material_models", [{}])
    if "material_ids" in obj and dim == 3:
        ids = meshio.read_dmat(res(obj["material_ids"])).reshape(-1)[:nT].astype(np.int64)
    else:
        idx = 0 if "material_ids" in obj else int(obj.get("material_index", 0))
        ids = np.full(nT, idx, dtype=np.int64)
    E = np.array([m.get("youngs_modulus", 1e6) for m in mats])[ids]
    nu = np.array([m.get("poissons_ratio", 0.45) for m in mats])[ids]
    mu, la = enu_to_lame(E, nu)
    mat = np.array([MATERIALS[m.get("energy", "Stable-Neohookean")] for m in mats],
                   dtype=np.int32)[ids]
    bc = obj.get("boundary_condition", {})
    fixed = bc_fixed(bc.get("type", "null"), V, bc.get("axis", 0), bc.get("ratio", 0.1))
    script = None
    if bc.get("type", "null") in ("stretch", "twist", "bend", "scale"):
        script = dict(type=bc["type"], axis=bc.get("axis", 0), ratio=bc.get("ratio", 0.1),
                      velocity=bc.get("velocity", 0.1))
    v0 = x0 = None
    if "initial_velocity" in obj:
        v0 = np.tile(np.asarray(obj["initial_velocity"], dtype=np.float64)[:dim], (V.shape[0], 1))
    st = d.get("initial_state", {})
    if "x_path" in st:
        x0 = meshio.read_dmat(res(st["x_path"]))
        if x0.shape != V.shape:
            raise ValueError("initial_state x_path has the wrong shape")
    if "v_path" in st:
        v0 = meshio.read_dmat(res(st["v_path"]))
        if v0.shape != V.shape:
            raise ValueError("initial_state v_path has the wrong shape")
    g = d.get("body_force", [0, -9.8, 0])
    return Scene(V=V, T=T, mu=mu, la=la, mat=mat, fixed=fixed,
                 density=mats[0].get("density", 1000.0), dt=d.get("dt", 0.034),
                 gravity=tuple(float(np.float32(x)) for x in g),
                 bdf_order=int(d.get("time_integrator", "BDF1")[3:]),
                 mixed="mixed-stretch" in d.get("mixed_variables", []),
                 solver=d.get("linear_solver", "eigen-llt"),
                 itr_tol=d.get("iterative_solver_tolerance", 1e-4),
                 max_iters=d.get("max_iterative_solver_iters", 500),
                 max_newton=d.get("max_newton_iterations", 5),
                 newton_tol=d.get("newton_tol", 1e-10),
                 ls_iters=d.get("max_linesearch_iterations", 20), bc=script, v0=v0, x0=x0,
                 name=os.path.basename(path))


# --- the benchmark configurations --------------------------------------------

def config_c1():
    """2D plain method, 40x10 rectangle of 800 triangles, SNH E=1e5 nu=0.45,
    BDF1 dt=1/30, gravity, pinned end edge, Newton to 1e-8, direct solve."""
    return rect2d_scene(40, 10, 0.1, E=1e5, nu=0.45, solver="eigen-llt",
                        max_newton=100, newton_tol=1e-8, dt=1.0 / 30.0)


def config_c2(nx=40, ny=10, nz=10, seed=0):
    """3D plain method FP64: 40x10x10x6 jittered beam (24,000 tets), SNH E=1e6,
    nu=0.45, one end clamped, gravity, BDF1, Jacobi PCG to rel-res 1e-10."""
    return beam_scene(nx, ny, nz, 0.1, 0.1, seed, E=1e6, nu=0.45, mixed=False,
                      solver="eigen-pcg-diag", itr_tol=1e-10, max_iters=20000,
                      dt=1.0 / 30.0)


def config_c3(nx=100, ny=40, nz=40, seed=0, **kw):
    """Mixed method at ~1M tets: 100x40x40x6 jittered beam (960,000 tets),
    per-tet E log-uniform in [1e5, 1e6], nu=0.45, SNH."""
    kw.setdefault("itr_tol", 1e-4)
    kw.setdefault("max_iters", 500)
    return beam_scene(nx, ny, nz, 0.1, 0.1, seed, E_range=(1e5, 1e6), nu=0.45,
                      mixed=True, solver="eigen-pcg-diag", dt=1.0 / 30.0, **kw)


def config_c5(nx=130, ny=50, nz=50, seed=0, **kw):
    """Stiff, ill-conditioned mixed beam: 130x50x50x6 jittered beam (1,950,000
    tets), per-tet E log-uniform in [1e4, 1e8], nu = 0.49, ARAP on one side
    of a seeded random plane and Stable Neo-Hookean on the other (per-element
    materials, as with the reference's material_ids, tet_mesh.h:11-15).

    The plane's normal is uniform on the sphere (stream 3 of the seed) and it
    passes through a uniform point of the beam's box (stream 4); a tet is ARAP
    when its centroid lies on the positive side."""
    kw.setdefault("itr_tol", 1e-4)
    kw.setdefault("max_iters", 500)
    h = 0.1
    sc = beam_scene(nx, ny, nz, h, 0.1, seed, E_range=(1e4, 1e8), nu=0.49, mixed=True,
                    solver="eigen-pcg-diag", dt=1.0 / 30.0, **kw)
    u = splitmix64_uniform(seed, 3, 2)
    z = 2.0 * u[0] - 1.0
    phi = 2.0 * math.pi * u[1]
    rxy = math.sqrt(max(0.0, 1.0 - z * z))
    normal = np.array([rxy * math.cos(phi), rxy * math.sin(phi), z])
    point = splitmix64_uniform(seed, 4, 3) * np.array([nx * h, ny * h, nz * h])
    centroid = sc.V[sc.T].mean(axis=1)
    arap = (centroid - point) @ normal > 0.0
    sc.mat = np.where(arap, MATERIALS["ARAP"], MATERIALS["Stable-Neohookean"]).astype(np.int32)
    sc.name = f"c5_{nx}x{ny}x{nz}"
    sc.extra = {"plane_normal": normal.tolist(), "plane_point": point.tolist(),
                "arap_fraction": float(arap.mean())}
    return sc


def c4_beam_sizes(count=256, seed=0, lo=2e3, hi=6e4):
    """Cell counts (nx, ny, nz = ny) of the batching config: 6 nx ny^2 drawn
    log-uniform in [lo, hi] (stream 5), the aspect nx / ny uniform in [2, 6]
    (stream 6), nx clipped to [10, 60]; ny >= 2."""
    t = np.exp(math.log(lo) + splitmix64_uniform(seed, 5, count) * (math.log(hi) - math.log(lo)))
    aspect = 2.0 + 4.0 * splitmix64_uniform(seed, 6, count)
    out = []
    for k in range(count):
        cells = t[k] / 6.0
        ny = max(2, int(round((cells / aspect[k]) ** (1.0 / 3.0))))
        nx = int(min(60, max(10, round(cells / (ny * ny)))))
        out.append((nx, ny, ny))
    return out


def config_c4(count=256, seed=0, **kw):
    """Batching config: `count` independent mixed-stretch beams whose tet
    counts are log-uniform in [2e3, 6e4], each with its own E log-uniform in
    [1e4, 1e7] (stream 7) and nu uniform in [0.3, 0.48] (stream 8); SNH, BDF1
    dt = 1/30, CG tolerance 1e-4 / 500 iterations. Returns a list of Scenes
    (beam k uses seed + k for its jitter)."""
    kw.setdefault("itr_tol", 1e-4)
    kw.setdefault("max_iters", 500)
    sizes = c4_beam_sizes(count, seed)
    E = np.exp(math.log(1e4) + splitmix64_uniform(seed, 7, count) * (math.log(1e7) - math.log(1e4)))
    nu = 0.3 + 0.18 * splitmix64_uniform(seed, 8, count)
    scenes_ = []
    for k, (nx, ny, nz) in enumerate(sizes):
        sc = beam_scene(nx, ny, nz, 0.1, 0.1, seed + k, E=float(E[k]), nu=float(nu[k]), mixed=True,
                        solver="eigen-pcg-diag", dt=1.0 / 30.0, **kw)
        sc.name = f"c4_{k}_{nx}x{ny}x{nz}"
        scenes_.append(sc)
    return scenes_
