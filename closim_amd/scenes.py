"""Seeded benchmark scenes (the five workloads the cloth step is measured on).

Every scene is built with the reference's own construction sequence:
buildGrid -> preprocess -> createScene (mesh.ts, scene.ts), then perturbations
are applied to `scene.positions` only, so hinge rest angles stay pi and Dm^-1
comes from the flat uv layout.  Randomness is mulberry32 with the documented
integer seeds (the PRNG the reference's tests use, tests/membrane-hvp.test.ts:10-18).
"""
import math

import numpy as np

from .mesh import build_grid, preprocess, sphere_mesh
from .scene import (ClothMaterial, ContactParams, ContactSystem, create_scene)

H = 1.0 / 60.0


def mulberry32(seed):
    """mulberry32 (tests/membrane-hvp.test.ts:10-18): returns a float in [0,1)."""
    a = seed & 0xFFFFFFFF

    def imul(x, y):
        return ((x & 0xFFFFFFFF) * (y & 0xFFFFFFFF)) & 0xFFFFFFFF

    def nxt():
        nonlocal a
        a = (a + 0x6D2B79F5) & 0xFFFFFFFF
        t = imul(a ^ (a >> 15), 1 | a)
        t = ((t + imul(t ^ (t >> 7), 61 | t)) & 0xFFFFFFFF) ^ t
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0
    return nxt


def mulberry32_array(seed, count):
    r = mulberry32(seed)
    return np.array([r() for _ in range(count)], np.float64)


# Material shared by the strip-pull and drape workloads.
STRIP_MATERIAL = ClothMaterial(arealDensityKgM2=0.2, thickness=5e-4, stretchWarp=1e6, stretchWeft=5e5,
                               stretchCoupling=0.0, shear=2e5, bendWarp=1e-5, bendWeft=1e-5, damping=0.001)


def strip_pull(nx=19, ny=99, width=0.2, length=1.0):
    """Strip-pull: 20x100-vertex sheet (0.2 m x 1.0 m) hung vertically
    (positions (x, -z, 0), uv (x, z)), top row pinned, bottom row pinned and
    driven 5 cm downward over 60 steps (`strip_pull_targets`)."""
    pos, uv, idx = build_grid(nx, ny, width, length)
    P = pos.reshape(-1, 3)
    vert = np.stack([P[:, 0], -P[:, 2], np.zeros_like(P[:, 0])], 1).astype(np.float32).reshape(-1)
    mesh = preprocess(vert, uv, idx, STRIP_MATERIAL.arealDensityKgM2)
    scene = create_scene(mesh, STRIP_MATERIAL, (0.0, -9.81, 0.0))
    top = [i for i in range(nx + 1)]
    bottom = [ny * (nx + 1) + i for i in range(nx + 1)]
    for i in top + bottom:
        scene.pinned[i] = tuple(float(v) for v in scene.positions[i * 3:i * 3 + 3])
    scene.meta = dict(bottom=bottom, bottom_y0=[scene.positions[i * 3 + 1] for i in bottom],
                      steps=60, grad_tol=1e-6, name="strip_pull_20x100")
    return scene


def strip_pull_targets(scene, step):
    """Bottom-row pin targets before step `step` (0-based): y0 - 0.05*(step+1)/60."""
    out = []
    for i, y0 in zip(scene.meta["bottom"], scene.meta["bottom_y0"]):
        out.append((i, (float(scene.positions[i * 3]), y0 - 0.05 * (step + 1) / 60.0, 0.0)))
    return out


def sphere_drape(nx=255, seed=0x5EED01, d_hat=1e-3, mu=0.3, sphere_y=0.2):
    """Sphere + floor drape: (nx+1)^2 sheet, 1 m side at y = 0.5 with seeded
    +-1e-4 m jitter per coordinate, static sphere r = 0.25 centred
    (0.5, sphere_y, 0.5) at 64x32 resolution, floor y = 0.

    The sphere centre sits at y = 0.2 (top 5 cm below the sheet): a sphere
    resting on the floor would touch the sheet at y = 0.5 and the jitter
    would start vertices inside d_min, where every line-search trial of the
    reference's checkTrial is rejected and the sheet never moves."""
    pos, uv, idx = build_grid(nx, nx, 1.0, 1.0)
    P = pos.reshape(-1, 3).copy()
    P[:, 1] = 0.5
    mesh = preprocess(P.reshape(-1), uv, idx, STRIP_MATERIAL.arealDensityKgM2)
    scene = create_scene(mesh, STRIP_MATERIAL, (0.0, -9.81, 0.0))
    jit = mulberry32_array(seed, mesh.count * 3)
    scene.positions += (jit * 2 - 1) * 1e-4
    c = ContactSystem(ContactParams(dHatM=d_hat, kappaJ=50.0, dMinM=1e-4, frictionMu=mu, frictionEpsM=1e-4))
    c.setFloor(0.0)
    sp, si = sphere_mesh(0.25, 64, 32, center=(0.5, sphere_y, 0.5))
    c.setStaticMesh(sp, si)
    scene.contact = c
    scene.meta = dict(steps=120, name=f"sphere_floor_drape_{nx + 1}x{nx + 1}")
    return scene


def crumple_fold(nx=511, side=1.5, seed=0xC2F01D, y0=0.06, amp=0.05, d_hat=5e-4, mu=0.4):
    """Self-contact fold: (nx+1)^2 sheet of side 1.5 m, initial height
    y = y0 + sum_k a_k sin(kx_k x + kz_k z + phi_k) with K in {3..6} seeded
    waves whose |a_k| sum to `amp`, wave numbers in [4, 24] rad/m, dropped on
    the floor y = 0."""
    pos, uv, idx = build_grid(nx, nx, side, side)
    mesh = preprocess(pos, uv, idx, STRIP_MATERIAL.arealDensityKgM2)
    scene = create_scene(mesh, STRIP_MATERIAL, (0.0, -9.81, 0.0))
    r = mulberry32(seed)
    K = 3 + int(r() * 4)
    w = np.array([0.5 + r() for _ in range(K)])
    a = amp * w / w.sum()
    kx = np.array([(4 + 20 * r()) * (1 if r() < 0.5 else -1) for _ in range(K)])
    kz = np.array([(4 + 20 * r()) * (1 if r() < 0.5 else -1) for _ in range(K)])
    ph = np.array([2 * math.pi * r() for _ in range(K)])
    X = scene.positions.reshape(-1, 3)
    y = np.full(mesh.count, y0)
    for k in range(K):
        y = y + a[k] * np.sin(kx[k] * X[:, 0] + kz[k] * X[:, 2] + ph[k])
    X[:, 1] = y
    c = ContactSystem(ContactParams(dHatM=d_hat, kappaJ=50.0, dMinM=1e-4, frictionMu=mu, frictionEpsM=1e-4))
    c.setFloor(0.0)
    scene.contact = c
    scene.meta = dict(steps=200, name=f"self_contact_fold_{nx + 1}x{nx + 1}", waves=K)
    return scene


def floor_patch(nx=6, hover=0.05, mu=0.3, material=None):
    """floor-contact.test.ts:11-49 scene: small patch 5 cm above the floor."""
    from .scene import DEFAULT_MATERIAL, DEFAULT_CONTACT_PARAMS
    from dataclasses import replace
    pos, uv, idx = build_grid(nx, nx, 0.1, 0.1)
    P = pos.reshape(-1, 3).copy()
    P[:, 1] += hover
    mesh = preprocess(P.reshape(-1), uv, idx, 0.15)
    scene = create_scene(mesh, material or DEFAULT_MATERIAL, (0.0, -9.81, 0.0))
    c = ContactSystem(replace(DEFAULT_CONTACT_PARAMS, frictionMu=mu))
    c.setFloor(0.0)
    scene.contact = c
    return scene


def offset_mesh(nx, ny, w, h, ox, oy, oz):
    """offsetMesh() tests/helpers.ts:6-17."""
    pos, uv, idx = build_grid(nx, ny, w, h)
    P = pos.reshape(-1, 3).astype(np.float64)
    P[:, 0] = (P[:, 0] + ox)
    P[:, 1] = (P[:, 1] + oy)
    P[:, 2] = (P[:, 2] + oz)
    return P.astype(np.float32).reshape(-1), uv, idx


def merge_parts(parts):
    """mergeParts() tests/helpers.ts:19-36."""
    pos = np.concatenate([p[0] for p in parts]).astype(np.float32)
    uv = np.concatenate([p[1] for p in parts]).astype(np.float32)
    idx = []
    off = 0
    for p in parts:
        idx.append(p[2].astype(np.int64) + off)
        off += len(p[0]) // 3
    return pos, uv, np.concatenate(idx).astype(np.uint32)


def head_on_patches():
    """self-contact.test.ts:11-38: two 5x5 patches flying at each other (no gravity)."""
    from .scene import DEFAULT_MATERIAL, DEFAULT_CONTACT_PARAMS
    from dataclasses import replace
    A = offset_mesh(4, 4, 0.08, 0.08, -0.07, 0.02, 0)
    B = offset_mesh(4, 4, 0.08, 0.08, 0.07, 0.02, 0)
    pos, uv, idx = merge_parts([A, B])
    mesh = preprocess(pos, uv, idx, 0.15)
    scene = create_scene(mesh, DEFAULT_MATERIAL, (0.0, 0.0, 0.0))
    nA = len(A[0]) // 3
    V = scene.velocities.reshape(-1, 3)
    V[:nA, 0] = 0.3
    V[nA:, 0] = -0.3
    scene.contact = ContactSystem(replace(DEFAULT_CONTACT_PARAMS, frictionMu=0.0))
    return scene


def static_plate_patch(mu=0.3):
    """static-collider.test.ts:35-69: patch above a 2-triangle static plate."""
    from .scene import DEFAULT_MATERIAL, DEFAULT_CONTACT_PARAMS
    from dataclasses import replace
    g = offset_mesh(4, 4, 0.08, 0.08, 0.01, 0.08, 0.01)
    mesh = preprocess(*g, 0.15)
    scene = create_scene(mesh, DEFAULT_MATERIAL, (0.0, -9.81, 0.0))
    sp = np.array([-0.05, 0.03, -0.05, 0.15, 0.03, -0.05, 0.15, 0.03, 0.15, -0.05, 0.03, 0.15], np.float32)
    si = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    c = ContactSystem(replace(DEFAULT_CONTACT_PARAMS, frictionMu=mu))
    c.setStaticMesh(sp, si)
    scene.contact = c
    return scene


def small_strip():
    """smoke.test.ts:16-48: 13x7 strip, left column pinned, DEFAULT_MATERIAL."""
    from .scene import DEFAULT_MATERIAL, pin_column
    pos, uv, idx = build_grid(12, 6, 0.2, 0.1)
    mesh = preprocess(pos, uv, idx, 0.15)
    scene = create_scene(mesh, DEFAULT_MATERIAL, (0.0, -9.81, 0.0))
    pin_column(scene, lambda x, y, z, i: x < 1e-9)
    return scene


def batch_drapes(count=64, seed=0xB47C4, d_hat=1e-3, mu=0.3):
    """The batched uneven drape: `count` independent sheets, each with
    (nx+1)(ny+1) ~ logU[1e3, 1e5] vertices and aspect nx/ny ~ U[0.5, 2] over
    1 m^2 (side sqrt(aspect) x 1/sqrt(aspect)) at y = 0.5 with seeded +-1e-4 m
    jitter, material moduli drawn log-uniformly (C00, C11 in [1e5, 1e7], G in
    [5e4, 5e6], k in [1e-6, 1e-4], C01 = 0; density, thickness and damping of
    the strip material), each over its own static sphere (r ~ U[0.1, 0.4],
    top 5 cm below the sheet, 32x16 segments) and a floor at y = 0."""
    from dataclasses import replace
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n_target = float(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
        aspect = float(rng.uniform(0.5, 2.0))
        ny = max(2, int(round(np.sqrt(n_target / aspect))) - 1)
        nx = max(2, int(round(aspect * (ny + 1))) - 1)
        w, l = float(np.sqrt(aspect)), float(1.0 / np.sqrt(aspect))
        c00, c11 = (float(np.exp(rng.uniform(np.log(1e5), np.log(1e7)))) for _ in range(2))
        g = float(np.exp(rng.uniform(np.log(5e4), np.log(5e6))))
        k = float(np.exp(rng.uniform(np.log(1e-6), np.log(1e-4))))
        r = float(rng.uniform(0.1, 0.4))
        mat = replace(STRIP_MATERIAL, stretchWarp=c00, stretchWeft=c11, stretchCoupling=0.0, shear=g,
                      bendWarp=k, bendWeft=k)
        pos, uv, idx = build_grid(nx, ny, w, l)
        P = pos.reshape(-1, 3).copy()
        P[:, 1] = 0.5
        mesh = preprocess(P.reshape(-1), uv, idx, mat.arealDensityKgM2)
        scene = create_scene(mesh, mat, (0.0, -9.81, 0.0))
        jit = mulberry32_array((seed + 7919 * i) & 0xFFFFFFFF, mesh.count * 3)
        scene.positions += (jit * 2 - 1) * 1e-4
        c = ContactSystem(ContactParams(dHatM=d_hat, kappaJ=50.0, dMinM=1e-4, frictionMu=mu, frictionEpsM=1e-4))
        c.setFloor(0.0)
        sp, si = sphere_mesh(r, 32, 16, center=(0.5 * w, 0.45 - r, 0.5 * l))
        c.setStaticMesh(sp, si)
        scene.contact = c
        scene.meta = dict(name=f"batch_drape_{i}_{nx + 1}x{ny + 1}", radius=r)
        out.append(scene)
    return out

