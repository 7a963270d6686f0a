"""Small contact scenes shared by the contact parity tests (CPU-sized)."""
import numpy as np

from closim_amd import mesh as M, scenes as S
from closim_amd.scene import (DEFAULT_CONTACT_PARAMS, DEFAULT_MATERIAL, ContactSystem, create_scene)
from dataclasses import replace


def stacked_patches(gap=1.5e-3, nx=6, mu=0.3):
    """Two parallel patches `gap` apart (self contact: VT + EE actives)."""
    A = S.offset_mesh(nx, nx, 0.06, 0.06, 0.0, 0.02, 0.0)
    B = S.offset_mesh(nx, nx, 0.06, 0.06, 0.003, 0.02 + gap, 0.002)
    pos, uv, idx = S.merge_parts([A, B])
    mesh = M.preprocess(pos, uv, idx, 0.15)
    scene = create_scene(mesh, DEFAULT_MATERIAL)
    scene.contact = ContactSystem(replace(DEFAULT_CONTACT_PARAMS, frictionMu=mu))
    return scene


def plate_close(height=0.0315):
    """Patch just above the static plate of static-collider.test.ts."""
    scene = S.static_plate_patch()
    X = scene.positions.reshape(-1, 3)
    X[:, 1] = height
    return scene


def floor_close(height=8e-4):
    scene = S.floor_patch()
    X = scene.positions.reshape(-1, 3)
    X[:, 1] = height
    return scene


def crumpled_sheet(nx=10, seed=3):
    """A crumpled sheet folded onto itself with floor (VT, EE and floor)."""
    pos, uv, idx = M.build_grid(nx, nx, 0.1, 0.1)
    mesh = M.preprocess(pos, uv, idx, 0.15)
    scene = create_scene(mesh, DEFAULT_MATERIAL)
    X = scene.positions.reshape(-1, 3)
    # fold along x = 0.05: the right half mirrored above the left half
    right = X[:, 0] > 0.05
    X[right, 0] = 0.1 - X[right, 0]
    X[right, 1] = 0.0025 + 0.02 * (X[right, 0] - 0.05) ** 2
    X[~right, 1] = 0.0012
    rng = np.random.default_rng(seed)
    X += rng.normal(0, 2e-4, X.shape)
    c = ContactSystem(replace(DEFAULT_CONTACT_PARAMS, frictionMu=0.3))
    c.setFloor(0.0)
    scene.contact = c
    return scene


def stacked_large():
    """Two stacked 33x33-cell patches (4,356 triangles): a hierarchy five
    levels deep, so every refit level and multi-wave traversal is exercised."""
    return stacked_patches(gap=1.2e-3, nx=33)


def sphere_touch():
    """48x48-vertex drape sheet 0.5 mm above the top of the 64x32 sphere
    (4,096 static triangles): cloth/static candidates through both hierarchies."""
    return S.sphere_drape(47, sphere_y=0.2495)


ALL = {"stacked": stacked_patches, "plate": plate_close, "floor": floor_close, "crumpled": crumpled_sheet,
       "stacked_large": stacked_large, "sphere_touch": sphere_touch}
