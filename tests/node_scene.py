"""Scenes written as JSON for the Node-API addon's test scripts (tests/node/),
which read them back into the reference's ClothScene shape."""
import json

import numpy as np


def _fl(a):
    return [float(v) for v in np.asarray(a).ravel()]


def write_scene(scene, path):
    m = scene.mesh
    hv = np.asarray(m.hinge_verts, np.int64).reshape(-1, 4)
    hinges = [[int(a), int(b), int(c), int(d), float(r), float(l), float(s)]
              for (a, b, c, d), r, l, s in zip(hv, m.hinge_rest, m.hinge_len, m.hinge_area)]
    c = scene.contact
    contact = None
    if c is not None:
        p = c.params
        contact = dict(params=dict(dHatM=p.dHatM, kappaJ=p.kappaJ, dMinM=p.dMinM, frictionMu=p.frictionMu,
                                   frictionEpsM=p.frictionEpsM),
                       floorY=None if c.floorY is None else float(c.floorY),
                       staticPos=None if c.staticPos is None else _fl(c.staticPos),
                       staticIdx=None if c.staticIdx is None else [int(v) for v in np.asarray(c.staticIdx).ravel()])
    mat = scene.material
    doc = dict(
        mesh=dict(count=int(m.count), triCount=int(m.tri_count), positions=_fl(m.positions),
                  restPositions=_fl(m.rest_positions), uv=_fl(m.uv),
                  indices=[int(v) for v in np.asarray(m.indices).ravel()], invDm=_fl(m.inv_dm), areas=_fl(m.areas),
                  masses=_fl(m.masses), hinges=hinges),
        material={k: float(getattr(mat, k)) for k in ("arealDensityKgM2", "thickness", "stretchWarp", "stretchWeft",
                                                      "stretchCoupling", "shear", "bendWarp", "bendWeft", "damping")},
        gravity=[float(g) for g in scene.gravity],
        pinned=[[int(i), float(p[0]), float(p[1]), float(p[2])] for i, p in scene.pinned.items()],
        positions=_fl(scene.positions), velocities=_fl(scene.velocities), contact=contact)
    with open(path, "w") as f:
        json.dump(doc, f)
