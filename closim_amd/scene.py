"""ClothScene / ClothMaterial / ContactSystem configuration.

Host mirror of closim's src/physics/scene.ts, src/physics/types.ts and the
configuration half of src/collision/contact-assembly.ts.  A `ClothScene` is
what `ClothSolver.initialize(scene)` receives; `to_desc` packs it into the C
ABI struct (include/cloth_hip.h) without copying the arrays.
"""
import ctypes as C
from dataclasses import dataclass, field, replace
from typing import Dict, Optional, Tuple

import numpy as np

from . import _abi
from .mesh import ClothMeshData


@dataclass
class ClothMaterial:
    """ClothMaterial (physics/types.ts:1-14)."""
    arealDensityKgM2: float
    thickness: float
    stretchWarp: float
    stretchWeft: float
    stretchCoupling: float
    shear: float
    bendWarp: float
    bendWeft: float
    damping: float

    def to_c(self):
        return _abi.CsMaterial(self.arealDensityKgM2, self.thickness, self.stretchWarp,
                               self.stretchWeft, self.stretchCoupling, self.shear,
                               self.bendWarp, self.bendWeft, self.damping)


# DEFAULT_MATERIAL (physics/types.ts:16-26)
DEFAULT_MATERIAL = ClothMaterial(0.15, 0.001, 20000, 20000, 0, 5000, 1e-5, 1e-5, 0.001)


@dataclass
class ContactParams:
    """ContactParams (collision/types.ts:27-39)."""
    dHatM: float = 0.002
    kappaJ: float = 50.0
    dMinM: float = 1e-4
    frictionMu: float = 0.3
    frictionEpsM: float = 1e-4

    def to_c(self):
        return _abi.CsContactParams(self.dHatM, self.kappaJ, self.dMinM, self.frictionMu, self.frictionEpsM)


# DEFAULT_CONTACT_PARAMS (collision/types.ts:41-47)
DEFAULT_CONTACT_PARAMS = ContactParams()


class ContactSystem:
    """Configuration of the reference ContactSystem (contact-assembly.ts:46-73):
    params, floor and static collider mesh.  The frozen active set, lagged
    friction data and diagnostics live on the device inside the solver."""

    def __init__(self, params: ContactParams, cloth_indices=None):
        self.params = replace(params)
        self.floorY: Optional[float] = None
        self.staticPos: Optional[np.ndarray] = None
        self.staticIdx: Optional[np.ndarray] = None

    def setFloor(self, y):
        self.floorY = None if y is None else float(y)

    def setStaticMesh(self, pos, idx):
        self.staticPos = np.ascontiguousarray(pos, np.float32).reshape(-1)
        self.staticIdx = np.ascontiguousarray(idx, np.uint32).reshape(-1)


@dataclass
class ClothScene:
    """ClothScene (physics/scene.ts:5-14)."""
    mesh: ClothMeshData
    material: ClothMaterial
    gravity: Tuple[float, float, float]
    pinned: Dict[int, Tuple[float, float, float]]
    positions: np.ndarray
    velocities: np.ndarray
    contact: Optional[ContactSystem] = None


def create_scene(mesh: ClothMeshData, material: ClothMaterial, gravity=(0.0, -9.81, 0.0)) -> ClothScene:
    """createScene() scene.ts:16-26."""
    return ClothScene(mesh=mesh, material=replace(material), gravity=tuple(gravity), pinned={},
                      positions=mesh.positions.astype(np.float64),
                      velocities=np.zeros(mesh.count * 3, np.float64))


def pin_column(scene: ClothScene, predicate):
    """pinColumn() scene.ts:28-36."""
    P = scene.positions.reshape(-1, 3)
    for i in range(scene.mesh.count):
        x, y, z = (float(v) for v in P[i])
        if predicate(x, y, z, i):
            scene.pinned[i] = (x, y, z)


def enforce_pins(scene: ClothScene, x: np.ndarray):
    """enforcePins() scene.ts:38-44."""
    for i, p in scene.pinned.items():
        x[i * 3:i * 3 + 3] = p


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None and a.size else None


def to_desc(scene: ClothScene):
    """Pack a scene into CsSceneDesc.  Returns (desc, keepalive)."""
    m = scene.mesh
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a

    pos = arr(scene.positions, np.float64)
    vel = arr(scene.velocities, np.float64)
    idx = arr(m.indices, np.uint32)
    inv = arr(m.inv_dm, np.float32)
    areas = arr(m.areas, np.float32)
    masses = arr(m.masses, np.float32)
    rest = arr(m.rest_positions, np.float32)
    hv = arr(m.hinge_verts.reshape(-1), np.uint32)
    hr = arr(m.hinge_rest, np.float64)
    hl = arr(m.hinge_len, np.float64)
    ha = arr(m.hinge_area, np.float64)
    pin_ids = arr(np.array(list(scene.pinned.keys()), np.uint32), np.uint32)
    pin_pos = arr(np.array(list(scene.pinned.values()), np.float64).reshape(-1), np.float64)
    d = _abi.CsSceneDesc()
    d.n_verts, d.n_tris, d.n_hinges = m.count, m.tri_count, m.hinge_count
    d.positions = _ptr(pos, C.c_double)
    d.velocities = _ptr(vel, C.c_double)
    d.indices = _ptr(idx, C.c_uint32)
    d.inv_dm = _ptr(inv, C.c_float)
    d.areas = _ptr(areas, C.c_float)
    d.masses = _ptr(masses, C.c_float)
    d.rest_positions = _ptr(rest, C.c_float)
    d.hinge_verts = _ptr(hv, C.c_uint32)
    d.hinge_rest_angle = _ptr(hr, C.c_double)
    d.hinge_edge_len = _ptr(hl, C.c_double)
    d.hinge_area_sum = _ptr(ha, C.c_double)
    d.material = scene.material.to_c()
    d.gravity = (C.c_double * 3)(*scene.gravity)
    d.n_pins = len(scene.pinned)
    d.pin_ids = _ptr(pin_ids, C.c_uint32)
    d.pin_pos = _ptr(pin_pos, C.c_double)
    c = scene.contact
    d.contact_enabled = 1 if c is not None else 0
    if c is not None:
        d.contact = c.params.to_c()
        d.has_floor = 1 if c.floorY is not None else 0
        d.floor_y = c.floorY if c.floorY is not None else 0.0
        if c.staticPos is not None and c.staticIdx is not None and len(c.staticIdx):
            sp = arr(c.staticPos, np.float32)
            si = arr(c.staticIdx, np.uint32)
            d.n_static_verts = len(sp) // 3
            d.n_static_tris = len(si) // 3
            d.static_pos = _ptr(sp, C.c_float)
            d.static_idx = _ptr(si, C.c_uint32)
    return d, keep
