"""Seeded synthetic scenes and radar poses for the workloads the backend is measured on.

All randomness is SplitMix64 (Steele, Lea & Flood 2014): the i-th output of a stream with
seed s is mix(s + (i+1)·0x9E3779B97F4A7C15), so whole streams are generated vectorised.
Geometry is assembled with float32 arithmetic in the order of TriangleMesh::addQuad /
addBox (libhussar/include/hussar/core/mesh.h:30-54), so the meshes equal what the
reference's mesh builders would produce from the same corner points.

Workloads:
* ``dihedral``      — two 10 cm square PEC plates at 90° (4 triangles), corner at the origin,
                      fold along +y, co-located Tx/Rx 1.5 m away on the bisector, yawed about
                      y (examples/src/1_dihedral.cpp:48-112 is the prototype).
* ``vehicle``       — random boxes and panels filling a 2 m × 1.5 m × 5 m box (x, y, z) placed
                      10–30 m ahead of the radar on +z; exactly ``triangles`` triangles.
* ``precision``     — flat plates and trihedrals at 5, 50 and 150 m over a random-height
                      ground grid.
* ``radar_array``   — 1 Tx plus ``channels`` receivers spaced λ/2 along the array's x axis.
"""
from __future__ import annotations

import numpy as np

from .types import (FrameConfig, GHz, MHz, NFAntenna, RFConfig, Scene, TriangleMesh,
                    SPEED_OF_LIGHT, rotation_y, us)

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Outputs start .. start+n-1 of the SplitMix64 stream with the given seed (uint64)."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniforms(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n, start) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _quads(a: np.ndarray, b: np.ndarray, c: np.ndarray):
    """Vectorised TriangleMesh::addQuad: float32 vertices (n,4,3) and local indices."""
    a = a.astype(np.float32)
    ab = a + b.astype(np.float32)
    v = np.stack([a, ab, ab + c.astype(np.float32), a + c.astype(np.float32)], axis=1)
    return v


def _assemble(quads: np.ndarray) -> TriangleMesh:
    n = quads.shape[0]
    v = quads.reshape(-1, 3)
    base = (np.arange(n, dtype=np.int64) * 4)[:, None]
    tri = np.concatenate([base + np.array([0, 1, 2]), base + np.array([0, 2, 3])], axis=1).reshape(-1, 3)
    return TriangleMesh(v, tri.astype(np.int32))


def _box_quads(lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    """Vectorised TriangleMesh::addBox (mesh.h:42-54): (n, 6, 4, 3)."""
    lo = lo.astype(np.float32)
    hi = hi.astype(np.float32)
    d = hi - lo
    z = np.zeros_like(d[:, 0])

    def vec(x, y, zz):
        return np.stack([x, y, zz], axis=1)

    q = [
        _quads(lo, vec(z, d[:, 1], z), vec(z, z, d[:, 2])),
        _quads(lo, vec(z, z, d[:, 2]), vec(d[:, 0], z, z)),
        _quads(lo, vec(d[:, 0], z, z), vec(z, d[:, 1], z)),
    ]
    d = lo - hi
    q += [
        _quads(hi, vec(z, d[:, 1], z), vec(d[:, 0], z, z)),
        _quads(hi, vec(z, z, d[:, 2]), vec(z, d[:, 1], z)),
        _quads(hi, vec(d[:, 0], z, z), vec(z, z, d[:, 2])),
    ]
    return np.stack(q, axis=1)


# --------------------------------------------------------------------------------------
# Dihedral corner reflector

FACING_NEG_X = np.array([[0, 0, -1], [0, -1, 0], [-1, 0, 0]], np.float64)  # 1_dihedral.cpp:96-100


def dihedral_mesh(size: float = 0.1) -> TriangleMesh:
    """Two zero-thickness size×size plates: x=0 (y,z ∈ [0,size]) and z=0 (x,y ∈ [0,size])."""
    m = TriangleMesh()
    s = np.float32(size)
    m.addQuad([0, 0, 0], [0, s, 0], [0, 0, s])
    m.addQuad([0, 0, 0], [s, 0, 0], [0, s, 0])
    return m


def dihedral_rf() -> RFConfig:
    """FMCW chirp of the dihedral sweep: 77 GHz, 60 MHz/µs, 10 MS/s, 256 samples (25.6 µs)."""
    return RFConfig(startFreq=77 * GHz, freqSlope=60 * MHz / us, adcRate=10 * MHz,
                    idleTime=100 * us, rampTime=25.6 * us, antennaDelay=0.0)


def dihedral_pose(yaw_deg: float, distance: float = 1.5, height: float = 0.05,
                  rf: RFConfig | None = None) -> Scene:
    """Co-located Tx/Rx at ``distance`` on the bisector, yawed by ``yaw_deg`` about +y; position
    and facing are both rotated by (yaw - 45°) as in 1_dihedral.cpp:90-112."""
    R = rotation_y(np.deg2rad(yaw_deg - 45.0))
    pos = R @ np.array([distance, height, 0.0])
    rot = R @ FACING_NEG_X
    ant = NFAntenna(pos.astype(np.float32), rot.astype(np.float32))
    return Scene(rfConfig=rf or dihedral_rf(), tx=ant, rx=ant)


def dihedral_sweep(step_deg: float = 1.0, lo: float = -45.0, hi: float = 45.0):
    n = int(round((hi - lo) / step_deg)) + 1
    return [lo + k * step_deg for k in range(n)]


# --------------------------------------------------------------------------------------
# Vehicle-like clutter of boxes and panels

def vehicle_mesh(seed: int = 1, triangles: int = 200_000, extent=(2.0, 1.5, 5.0),
                 range_lo: float = 10.0, range_hi: float = 30.0, edge_lo: float = 0.05,
                 edge_hi: float = 0.5):
    """Random boxes (12 triangles) and panels (2 triangles) inside an AABB of ``extent``
    (x width, y height, z length) standing on y = 0, centred at x = 0 and at a range drawn
    from U[range_lo, range_hi] on +z. Returns (mesh, centre)."""
    if triangles % 2:
        raise ValueError("triangle count must be even")
    head = uniforms(seed, 1)
    zc = range_lo + head[0] * (range_hi - range_lo)
    ext = np.asarray(extent, np.float64)
    lo_box = np.array([-ext[0] / 2, 0.0, zc - ext[2] / 2])
    # primitive types: box with probability 1/2 while at least 12 triangles remain
    n_max = triangles // 2
    u = uniforms(seed, 8 * n_max, start=1).reshape(n_max, 8)
    kinds = []
    remaining = triangles
    k = 0
    while remaining > 0:
        is_box = remaining >= 12 and u[k, 0] < 0.5
        kinds.append(is_box)
        remaining -= 12 if is_box else 2
        k += 1
    kinds = np.asarray(kinds, bool)
    u = u[:len(kinds)]
    quads = []
    # boxes
    ub = u[kinds]
    if len(ub):
        size = edge_lo + ub[:, 1:4] * (edge_hi - edge_lo)
        size = np.minimum(size, ext)
        lo = lo_box + ub[:, 4:7] * (ext - size)
        quads.append(_box_quads(lo, lo + size).reshape(-1, 4, 3))
    # panels
    up = u[~kinds]
    if len(up):
        a = edge_lo + up[:, 1] * (edge_hi - edge_lo)
        b = edge_lo + up[:, 2] * (edge_hi - edge_lo)
        cos_t = 2.0 * up[:, 3] - 1.0
        phi = 2.0 * np.pi * up[:, 4]
        sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t ** 2))
        nrm = np.stack([sin_t * np.cos(phi), cos_t, sin_t * np.sin(phi)], axis=1)
        helper = np.where(np.abs(nrm[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
        e1 = np.cross(nrm, helper)
        e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
        e2 = np.cross(nrm, e1)
        half = 0.5 * np.sqrt(a * a + b * b)[:, None]
        cen = lo_box + half + np.stack([up[:, 5], up[:, 6], up[:, 7]], axis=1) * np.maximum(ext - 2 * half, 0.0)
        corner = cen - 0.5 * a[:, None] * e1 - 0.5 * b[:, None] * e2
        quads.append(_quads(corner, a[:, None] * e1, b[:, None] * e2))
    # keep generation order: interleave by primitive index
    order_q = []
    bi = pi = 0
    box_q = quads[0].reshape(-1, 6, 4, 3) if len(ub) else None
    pan_q = quads[-1] if len(up) else None
    for is_box in kinds:
        if is_box:
            order_q.append(box_q[bi]); bi += 1
        else:
            order_q.append(pan_q[pi:pi + 1]); pi += 1
    mesh = _assemble(np.concatenate(order_q, axis=0))
    assert mesh.triangle_count() == triangles
    return mesh, np.array([0.0, ext[1] / 2, zc])


def _grid(origin, du, dv, nu, nv, rough, normal, seed):
    """nu x nv quads spanning origin + [0,1]du + [0,1]dv with vertices displaced along
    `normal` by U[-rough, rough] (shared vertices, 2 triangles per quad)."""
    uu, vv = np.meshgrid(np.linspace(0.0, 1.0, nu + 1), np.linspace(0.0, 1.0, nv + 1), indexing="ij")
    off = (uniforms(seed, (nu + 1) * (nv + 1)) * 2.0 - 1.0) * rough
    V = (np.asarray(origin)[None] + uu.reshape(-1, 1) * np.asarray(du)[None] + vv.reshape(-1, 1) * np.asarray(dv)[None]
         + off[:, None] * np.asarray(normal)[None])
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    v00 = (ii * (nv + 1) + jj).reshape(-1)
    v10, v01 = v00 + (nv + 1), v00 + 1
    v11 = v10 + 1
    T = np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)], axis=1).reshape(-1, 3)
    return V.astype(np.float32), T.astype(np.int32)


def urban_mesh(seed: int = 4, facade_nz: int = 1000, facade_ny: int = 300, ground_nx: int = 400,
               ground_nz: int = 1000, vehicles: int = 4, poles: int = 20):
    """A street canyon along +z: two rough facades (x = ±8 m, 15 m high, z 5..105 m), a rough
    street surface (y = 0), `vehicles` vehicle-like clutter blocks in lanes x = ±2 m at
    10..40 m moving along ±z at U[0, 20] m/s, and `poles` static poles. With the defaults:
    2,000,000 facade and ground triangles plus 8,000 vehicle and 240 pole triangles. Returns
    (mesh, per-triangle velocities float32 [T, 3])."""
    m = TriangleMesh()
    vel = []
    parts = [
        _grid([-8.0, 0.0, 5.0], [0.0, 0.0, 100.0], [0.0, 15.0, 0.0], facade_nz, facade_ny, 0.05, [1.0, 0, 0], seed),
        _grid([8.0, 0.0, 5.0], [0.0, 15.0, 0.0], [0.0, 0.0, 100.0], facade_ny, facade_nz, 0.05, [1.0, 0, 0], seed + 1),
        _grid([-8.0, 0.0, 5.0], [16.0, 0.0, 0.0], [0.0, 0.0, 100.0], ground_nx, ground_nz, 0.02, [0, 1.0, 0], seed + 2),
    ]
    for V, T in parts:
        m.append(V, T)
        vel.append(np.zeros((T.shape[0], 3), np.float32))
    u = uniforms(seed + 3, 2 * vehicles).reshape(vehicles, 2)
    for k in range(vehicles):
        car, _ = vehicle_mesh(seed=40 + seed + k, triangles=2000, extent=(1.8, 1.5, 4.5), range_lo=10.0, range_hi=40.0)
        V = car.vertexBuffer.copy()
        V[:, 0] += np.float32(-2.0 if k % 2 == 0 else 2.0)
        m.append(V, car.indexBuffer)
        speed = 20.0 * u[k, 0] * (1.0 if u[k, 1] < 0.5 else -1.0)
        vel.append(np.tile(np.float32([0.0, 0.0, speed]), (car.triangle_count(), 1)))
    if poles:
        zs = 7.5 + 5.0 * np.arange(poles // 2)
        lo = np.concatenate([np.stack([np.full_like(zs, -6.5), np.zeros_like(zs), zs], 1),
                             np.stack([np.full_like(zs, 6.35), np.zeros_like(zs), zs], 1)])
        pm = _assemble(_box_quads(lo, lo + np.array([0.15, 6.0, 0.15])).reshape(-1, 4, 3))
        m.append(pm.vertexBuffer, pm.indexBuffer)
        vel.append(np.zeros((pm.triangle_count(), 3), np.float32))
    return m, np.concatenate(vel)


def mimo_array(rf: RFConfig, tx: int = 3, rx: int = 4, position=(0.0, 1.0, 0.0)) -> Scene:
    """TDM-MIMO array looking down +z: rx receivers at λ/2 and tx transmitters at rx·λ/2
    along x, giving tx·rx virtual channels on a λ/2 grid."""
    base = radar_array(rf, channels=rx, position=position)
    lam = SPEED_OF_LIGHT / rf.startFreq
    p = np.asarray(position, np.float64)
    txs = [NFAntenna((p + k * rx * 0.5 * lam * np.array([1.0, 0, 0])).astype(np.float32), np.eye(3, dtype=np.float32))
           for k in range(tx)]
    return Scene(rfConfig=rf, tx=txs[0], rx=base.rx, rxs=base.rxs, txs=txs)


def radar_array(rf: RFConfig, channels: int = 4, position=(0.0, 0.75, 0.0),
                rotation: np.ndarray | None = None) -> Scene:
    """1 Tx at ``position`` and ``channels`` receivers at position + k·λ/2 along the array's
    local x axis (λ at the chirp start frequency); boresight = local +z."""
    R = np.eye(3) if rotation is None else np.asarray(rotation, np.float64)
    lam = SPEED_OF_LIGHT / rf.startFreq
    p = np.asarray(position, np.float64)
    tx = NFAntenna(p.astype(np.float32), R.astype(np.float32))
    rxs = [NFAntenna((p + k * 0.5 * lam * R[:, 0]).astype(np.float32), R.astype(np.float32))
           for k in range(channels)]
    return Scene(rfConfig=rf, tx=tx, rx=rxs[0], rxs=rxs)


def plate_mesh(size: float = 0.5, distance: float = 5.0) -> TriangleMesh:
    """A size×size plate in the plane z = distance facing a radar at the origin looking down
    +z (2 triangles)."""
    m = TriangleMesh()
    h = np.float32(size / 2)
    m.addQuad([-h, -h, np.float32(distance)], [np.float32(size), 0, 0], [0, np.float32(size), 0])
    return m


# --------------------------------------------------------------------------------------
# Precision study: plates and trihedrals at 5 / 50 / 150 m over random-height ground

def trihedral_quads(corner, size: float, R: np.ndarray) -> np.ndarray:
    """Three mutually orthogonal square plates sharing ``corner`` (6 triangles as 3 quads)."""
    c = np.asarray(corner, np.float64)
    ax = [R[:, k] * size for k in range(3)]
    qa = [(ax[0], ax[1]), (ax[1], ax[2]), (ax[2], ax[0])]
    return np.concatenate([_quads(c[None], a[None], b[None]) for a, b in qa], axis=0)


def precision_mesh(seed: int = 2, ground_nx: int = 500, ground_nz: int = 1000,
                   ground_size=(50.0, 200.0), rms: float = 0.05):
    """Targets on +z at 5, 50 and 150 m (a plate and a trihedral each) plus a ground grid of
    ground_nx × ground_nz quads (2 triangles each) with heights N(0, rms), y = -1.5 m."""
    quads = []
    for k, rng in enumerate((5.0, 50.0, 150.0)):
        # flat plate facing the radar (normal -z), 0.5 m
        s = 0.5
        quads.append(_quads(np.array([[-1.5 - s / 2, -s / 2, rng]]), np.array([[s, 0, 0]]),
                            np.array([[0, s, 0]])))
        # trihedral opening towards the radar
        R = np.array([[1, 0, 0], [0, 1, 0], [0, 0, -1]], np.float64)
        quads.append(trihedral_quads([1.5, -0.25, rng + 0.25], 0.3, R))
    targets = np.concatenate(quads, axis=0)
    mesh = _assemble(targets)
    # ground grid with shared vertices
    nx, nz = ground_nx, ground_nz
    xs = np.linspace(-ground_size[0] / 2, ground_size[0] / 2, nx + 1)
    zs = np.linspace(0.0, ground_size[1], nz + 1)
    g = uniforms(seed, 2 * (nx + 1) * (nz + 1)).reshape(-1, 2)
    # Box–Muller normal heights
    h = rms * np.sqrt(-2.0 * np.log(np.maximum(g[:, 0], 1e-300))) * np.cos(2 * np.pi * g[:, 1])
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    V = np.stack([X.reshape(-1), -1.5 + h, Z.reshape(-1)], axis=1).astype(np.float32)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    v00 = (ii * (nz + 1) + jj).reshape(-1)
    v10 = v00 + (nz + 1)
    v01 = v00 + 1
    v11 = v10 + 1
    T = np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)], axis=1).reshape(-1, 3)
    mesh.append(V, T.astype(np.int32))
    return mesh
