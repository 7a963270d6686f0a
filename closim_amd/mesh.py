"""Cloth mesh construction: host mirror of closim's src/mesh/mesh.ts.

`build_grid` and `preprocess` reproduce the reference's arrays bit for bit:
float32 storage of invDm/areas/masses (mesh.ts:3-14), lumped masses accumulated
in triangle order with a float32 store after every add (mesh.ts:82-85), hinges
in Map insertion order of the first occurrence of each interior edge
(mesh.ts:99-126) and V8's Math.hypot for edge lengths / dihedral normals.
"""
import math
from dataclasses import dataclass, field

import numpy as np


def js_hypot3(a, b, c):
    """Vectorised V8 Math.hypot: max-scaled Kahan sum of squares, sqrt * max."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c = np.asarray(c, np.float64)
    ab = [np.abs(a), np.abs(b), np.abs(c)]
    mx = np.maximum(np.maximum(ab[0], ab[1]), ab[2])
    safe = np.where(mx > 0, mx, 1.0)
    s = np.zeros_like(mx)
    comp = np.zeros_like(mx)
    for v in ab:
        n = v / safe
        summand = n * n - comp
        prelim = s + summand
        comp = (prelim - s) - summand
        s = prelim
    out = np.sqrt(s) * mx
    out = np.where(mx == 0, 0.0, out)
    nan = np.isnan(a) | np.isnan(b) | np.isnan(c)
    out = np.where(nan, np.nan, out)
    return np.where(np.isinf(mx), np.inf, out)


_PS = (1.66666666666666657415e-01, -3.25565818622400915405e-01, 2.01212532134862925881e-01,
       -4.00555345006794114027e-02, 7.91534994289814532176e-04, 3.47933107596021167570e-05)
_QS = (-2.40339491173441421878e+00, 2.02094576023350569471e+00, -6.88283971605453293030e-01,
       7.70381505559019352791e-02)


def _pq(z):
    p = z * (_PS[0] + z * (_PS[1] + z * (_PS[2] + z * (_PS[3] + z * (_PS[4] + z * _PS[5])))))
    q = 1.0 + z * (_QS[0] + z * (_QS[1] + z * (_QS[2] + z * _QS[3])))
    return p / q


def _acos(x):
    """Math.acos as V8 computes it (fdlibm e_acos.c), vectorised with the same
    IEEE operations as the device code, so rest angles are bit-identical."""
    x = np.ascontiguousarray(x, np.float64)
    u = x.view(np.uint64)
    hx = (u >> np.uint64(32)).astype(np.int64)
    hx = np.where(hx >= 2 ** 31, hx - 2 ** 32, hx)
    ix = hx & 0x7fffffff
    lo = (u & np.uint64(0xffffffff)).astype(np.int64)
    pi, pio2_hi, pio2_lo = 3.14159265358979311600e+00, 1.57079632679489655800e+00, 6.12323399573676603587e-17
    out = np.full_like(x, np.nan)
    with np.errstate(all="ignore"):
        one = (ix == 0x3ff00000) & (lo == 0)
        out[one & (hx > 0)] = 0.0
        out[one & (hx <= 0)] = pi + 2.0 * pio2_lo
        small = ix < 0x3fe00000
        tiny = small & (ix <= 0x3c600000)
        out[tiny] = pio2_hi + pio2_lo
        s1 = small & ~tiny
        xs = x[s1]
        r = _pq(xs * xs)
        out[s1] = pio2_hi - (xs - (pio2_lo - xs * r))
        mid = (ix >= 0x3fe00000) & (ix < 0x3ff00000)
        neg = mid & (hx < 0)
        z = (1.0 + x[neg]) * 0.5
        s = np.sqrt(z)
        w = _pq(z) * s - pio2_lo
        out[neg] = pi - 2.0 * (s + w)
        pos = mid & (hx >= 0)
        z = (1.0 - x[pos]) * 0.5
        s = np.sqrt(z)
        df = (s.view(np.uint64) & np.uint64(0xffffffff00000000)).view(np.float64)
        c = (z - df * df) / (s + df)
        w = _pq(z) * s + c
        out[pos] = 2.0 * (df + w)
    return out


def dihedral(x, v0, v1, v2, v3):
    """dihedral() mesh.ts:129-140 over arrays of hinge vertex ids."""
    x = np.asarray(x)
    X = x.reshape(-1, 3).astype(np.float64)
    e0 = X[v1] - X[v0]
    e1 = X[v2] - X[v0]
    e2 = X[v3] - X[v0]
    n1 = np.stack([e0[:, 1] * e1[:, 2] - e0[:, 2] * e1[:, 1],
                   e0[:, 2] * e1[:, 0] - e0[:, 0] * e1[:, 2],
                   e0[:, 0] * e1[:, 1] - e0[:, 1] * e1[:, 0]], 1)
    n2 = np.stack([e0[:, 1] * e2[:, 2] - e0[:, 2] * e2[:, 1],
                   e0[:, 2] * e2[:, 0] - e0[:, 0] * e2[:, 2],
                   e0[:, 0] * e2[:, 1] - e0[:, 1] * e2[:, 0]], 1)
    l1 = js_hypot3(n1[:, 0], n1[:, 1], n1[:, 2]) + 1e-30
    l2 = js_hypot3(n2[:, 0], n2[:, 1], n2[:, 2]) + 1e-30
    dot = n1[:, 0] * n2[:, 0] + n1[:, 1] * n2[:, 1] + n1[:, 2] * n2[:, 2]
    cos_t = np.minimum(1.0, np.maximum(-1.0, dot / (l1 * l2)))
    return _acos(cos_t)


def build_grid(nx, ny, w, h):
    """buildGrid() mesh.ts:17-48: flat XZ grid, uv = (x, z), tris (a,c,b),(b,c,d)."""
    i = np.arange(nx + 1, dtype=np.float64)
    j = np.arange(ny + 1, dtype=np.float64)
    xs = (i / nx) * w
    zs = (j / ny) * h
    X, Z = np.meshgrid(xs, zs)  # rows = j
    n = (nx + 1) * (ny + 1)
    pos = np.zeros((n, 3), np.float32)
    pos[:, 0] = X.ravel()
    pos[:, 2] = Z.ravel()
    uv = np.stack([X.ravel(), Z.ravel()], 1).astype(np.float32)
    jj, ii = np.meshgrid(np.arange(ny, dtype=np.uint32), np.arange(nx, dtype=np.uint32), indexing="ij")
    a = jj * (nx + 1) + ii
    b = a + 1
    c = a + (nx + 1)
    d = c + 1
    idx = np.stack([a, c, b, b, c, d], -1).reshape(-1).astype(np.uint32)
    return pos.reshape(-1), uv.reshape(-1), idx


@dataclass
class ClothMeshData:
    """ClothMeshData (mesh.ts:3-14)."""
    count: int
    tri_count: int
    positions: np.ndarray        # float32 3n (initial)
    rest_positions: np.ndarray   # float32 3n
    uv: np.ndarray               # float32 2n
    indices: np.ndarray          # uint32 3m
    inv_dm: np.ndarray           # float32 4m
    areas: np.ndarray            # float32 m
    masses: np.ndarray           # float32 n
    hinge_verts: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.uint32))
    hinge_rest: np.ndarray = field(default_factory=lambda: np.zeros(0))
    hinge_len: np.ndarray = field(default_factory=lambda: np.zeros(0))
    hinge_area: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def hinge_count(self):
        return len(self.hinge_rest)


def _build_hinges(indices, pos, areas):
    """buildHinges() mesh.ts:99-126."""
    tri = indices.reshape(-1, 3).astype(np.int64)
    m = len(tri)
    a = np.concatenate([tri[:, [0]], tri[:, [1]], tri[:, [2]]], 1)   # edge e: ids[e]
    b = np.concatenate([tri[:, [1]], tri[:, [2]], tri[:, [0]]], 1)   # ids[(e+1)%3]
    o = np.concatenate([tri[:, [2]], tri[:, [0]], tri[:, [1]]], 1)   # ids[(e+2)%3]
    a, b, o = a.reshape(-1), b.reshape(-1), o.reshape(-1)
    t_of = np.repeat(np.arange(m), 3)
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    key = lo * (int(indices.max()) + 1 if m else 1) + hi
    occ = np.arange(3 * m)
    order = np.lexsort((occ, key))          # by key, then occurrence
    ks = key[order]
    start = np.r_[True, ks[1:] != ks[:-1]]
    gid = np.cumsum(start) - 1
    counts = np.bincount(gid)
    first = order[start]                    # first occurrence of each key
    two = counts == 2
    g_first = first[two]
    starts_idx = np.flatnonzero(start)
    g_second = order[starts_idx[two] + 1]
    # Map insertion order = order of first occurrence
    srt = np.argsort(g_first, kind="stable")
    f = g_first[srt]
    s = g_second[srt]
    v0, v1, v2, v3 = a[f], b[f], o[f], o[s]
    rest = dihedral(pos, v0, v1, v2, v3)
    P = pos.reshape(-1, 3).astype(np.float64)
    d = P[v0] - P[v1]
    length = js_hypot3(d[:, 0], d[:, 1], d[:, 2])
    asum = areas[t_of[f]].astype(np.float64) + areas[t_of[s]].astype(np.float64)
    hv = np.stack([v0, v1, v2, v3], 1).astype(np.uint32)
    return hv, rest, length, asum


def preprocess(positions, uv, indices, areal_density):
    """preprocess() mesh.ts:50-97."""
    positions = np.asarray(positions, np.float32).reshape(-1)
    uv = np.asarray(uv, np.float32).reshape(-1)
    indices = np.asarray(indices, np.uint32).reshape(-1)
    n = len(positions) // 3
    m = len(indices) // 3
    tri = indices.reshape(-1, 3)
    UV = uv.reshape(-1, 2).astype(np.float64)
    u0, v0 = UV[tri[:, 0], 0], UV[tri[:, 0], 1]
    u1, v1 = UV[tri[:, 1], 0], UV[tri[:, 1], 1]
    u2, v2 = UV[tri[:, 2], 0], UV[tri[:, 2], 1]
    du1, dv1 = u1 - u0, v1 - v0
    du2, dv2 = u2 - u0, v2 - v0
    det = du1 * dv2 - du2 * dv1
    bad = np.flatnonzero(np.abs(det) < 1e-12)
    if len(bad):
        raise ValueError(f"degenerate triangle {int(bad[0])}")
    area = 0.5 * np.abs(det)
    areas = area.astype(np.float32)
    inv_dm = np.stack([dv2 / det, -du2 / det, -dv1 / det, du1 / det], 1).astype(np.float32).reshape(-1)
    tm = areal_density * area / 3
    # masses[i] += tm in triangle order with a float32 store after each add:
    # walk the per-vertex incidence lists (triangle order) slot by slot.
    corner_v = tri.reshape(-1).astype(np.int64)
    corner_tm = np.repeat(tm, 3)
    order = np.argsort(corner_v, kind="stable")
    cv = corner_v[order]
    ctm = corner_tm[order]
    starts = np.r_[0, np.flatnonzero(cv[1:] != cv[:-1]) + 1]
    seg = np.repeat(np.arange(len(starts)), np.diff(np.r_[starts, len(cv)]))
    slot = np.arange(len(cv)) - starts[seg]
    masses = np.zeros(n, np.float32)
    for k in range(int(slot.max()) + 1 if len(slot) else 0):
        sel = slot == k
        vv = cv[sel]
        masses[vv] = (masses[vv].astype(np.float64) + ctm[sel]).astype(np.float32)
    hv, rest, length, asum = _build_hinges(indices, positions, areas)
    return ClothMeshData(
        count=n, tri_count=m, positions=positions.copy(), rest_positions=positions.copy(),
        uv=uv.copy(), indices=indices.copy(), inv_dm=inv_dm, areas=areas, masses=masses,
        hinge_verts=hv, hinge_rest=rest, hinge_len=length, hinge_area=asum)


def sphere_mesh(radius, segments=12, rings=6, center=(0.0, 0.0, 0.0)):
    """Static sphere collider: sphereMesh() avatar-collider.ts:273-300 translated
    to `center` and stored as float32 (preprocessCollisionMesh :178-205; the
    sphere has no duplicate vertices or degenerate faces to weld/drop)."""
    if not (6 <= segments <= 64 and 3 <= rings <= 32):
        raise ValueError("sphere resolution out of range")
    r = float(radius)
    pos = [0.0, r, 0.0]
    for j in range(1, rings):
        theta = math.pi * j / rings
        for i in range(segments):
            phi = 2 * math.pi * i / segments
            pos += [r * math.sin(theta) * math.cos(phi), r * math.cos(theta), r * math.sin(theta) * math.sin(phi)]
    bottom = len(pos) // 3
    pos += [0.0, -r, 0.0]
    idx = []
    for i in range(segments):
        nxt = (i + 1) % segments
        idx += [0, 1 + i, 1 + nxt]
        for ring in range(rings - 2):
            a = 1 + ring * segments + i
            b = 1 + ring * segments + nxt
            c = a + segments
            d = b + segments
            idx += [a, c, b, b, c, d]
        last = 1 + (rings - 2) * segments
        idx += [last + i, bottom, last + nxt]
    P = np.array(pos, np.float64).reshape(-1, 3) + np.asarray(center, np.float64)
    P = P + 0.0  # canonical zero
    return P.astype(np.float32).reshape(-1), np.array(idx, np.uint32)
