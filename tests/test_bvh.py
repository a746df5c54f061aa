"""BVH invariants (product four-wide layout with quantized child boxes, and the
oracle's reference-shaped tree) and closest-hit equality with an O(N) brute-force
scan."""
import numpy as np
import pytest

import cudart_amd as ca
import oracle as O
from cudart_amd import workloads as W


def _tri_bounds(pos):
    return pos.min(axis=1), pos.max(axis=1)


def _check_product_layout(mesh):
    # a Mesh, or a tests.helpers.OwnedDesc
    a = mesh.arrays() if isinstance(mesh, ca.Mesh) else ca.scene_arrays(mesh.desc)
    s = ca.Scene(mesh if isinstance(mesh, ca.Mesh) else mesh.desc)
    info = s.bvh_info()
    nodes, recs, order = s.export_bvh()
    n = len(a.positions)
    assert sorted(order.tolist()) == list(range(n))
    # records: v1, e1 = v2 - v1, e2 = v3 - v1, original id and material
    p = a.positions[order]
    np.testing.assert_array_equal(recs[:, 0:3], p[:, 0])
    np.testing.assert_array_equal(recs[:, 4:7], p[:, 1] - p[:, 0])
    np.testing.assert_array_equal(recs[:, 8:11], p[:, 2] - p[:, 0])
    assert np.array_equal(recs[:, 3].view(np.uint32), order)
    assert np.array_equal(recs[:, 7].view(np.uint32), a.material_ids[order])
    u32 = nodes.view(np.uint32)
    words = nodes[:, 4:8].view(np.int32)
    lo, hi = _tri_bounds(a.positions)
    covered = np.zeros(n, np.int32)
    seen_nodes = np.zeros(len(nodes), np.int32)
    depth_max = 0
    stack = [(0, np.full(3, -np.inf), np.full(3, np.inf), 0)]
    while stack:
        k, plo, phi, depth = stack.pop()
        seen_nodes[k] += 1
        depth_max = max(depth_max, depth)
        origin = nodes[k, 0:3].astype(np.float32)
        ex = int(u32[k, 3])
        step = np.array([np.uint32(((ex >> (8 * ax)) & 0xFF) << 23).view(np.float32) for ax in range(3)], np.float32)
        assert (u32[k, 14:16] == 0).all() and (ex >> 24) == 0
        used = 0
        for c in range(4):
            w = int(words[k, c])
            q = np.array([(int(u32[k, 8 + j]) >> (8 * c)) & 0xFF for j in range(6)], np.float32)
            if w == 0x7FFFFFFF:  # empty slot: no planes, only after the used ones
                assert (q == 0).all()
                continue
            assert used == c, "empty slots come last"
            used += 1
            # decode as the kernel does: fmaf(q, 2^e, origin); q * 2^e is exact, so
            # the float32 sum is the single rounding of the fused operation
            clo = (q[0::2] * step + origin).astype(np.float32)
            chi = (q[1::2] * step + origin).astype(np.float32)
            assert (clo <= chi).all()
            if w >= 0:
                assert w < len(nodes)
                stack.append((w, clo, chi, depth + 1))
            else:
                u = w & 0xFFFFFFFF
                first, count = u & 0x0FFFFFFF, ((u >> 28) & 7) + 1
                assert count <= 8
                ids = order[first:first + count]
                assert (lo[ids] >= clo).all() and (hi[ids] <= chi).all(), "leaf box does not contain a triangle"
                assert (lo[ids] > clo).all() and (hi[ids] < chi).all(), "leaf box is not padded"
                covered[first:first + count] += 1
        assert used >= 1 and (used >= 2 or len(nodes) == 1), "an inner node holds at least two children"
        inner = [int(w) for w in words[k] if 0 <= w < 0x7FFFFFFF]
        if inner:
            assert inner == list(range(inner[0], inner[0] + len(inner))), "siblings are consecutive"
        # the children's decoded boxes lie inside the parent's decoded box for this node
        assert (origin >= plo).all()
    assert (covered == 1).all(), "every record in exactly one leaf"
    assert (seen_nodes == 1).all(), "every node reached once"
    assert depth_max == info.max_depth
    return info


def test_product_bvh_cornell():
    _check_product_layout(W.cornell().mesh)


def test_product_bvh_procedural_10k():
    info = _check_product_layout(ca.Mesh.generate("mixed", "tris=10000", 3))
    assert info.max_leaf_size <= 4


@pytest.mark.parametrize("kind", ["gauss", "clustered", "duplicates", "strips"])
def test_product_bvh_awkward_scenes(kind):
    """The collapsed four-wide layout on inputs that stress the builder and the
    collapse: tiny clustered triangles, many exact duplicates (coincident
    centroids), long thin strips, and sizes from 1 to a few thousand triangles."""
    from tests.helpers import OwnedDesc, material
    rng = np.random.default_rng(["gauss", "clustered", "duplicates", "strips"].index(kind))
    cam = ca.camera_look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), 40, 8, 8)
    for n in (1, 2, 3, 5, 9, 17, 64, 600, 2500):
        if kind == "gauss":
            p = rng.normal(size=(n, 3, 3))
        elif kind == "clustered":
            p = rng.normal(size=(n, 1, 3)) * 10 + rng.normal(size=(n, 3, 3)) * 1e-3
        elif kind == "duplicates":
            base = rng.normal(size=(1 + n // 50, 3, 3))
            p = base[rng.integers(0, len(base), n)]
        else:
            p = rng.normal(size=(n, 3, 3)) * np.array([100.0, 1e-4, 1.0])
        p = p.astype(np.float32)
        nrm = np.zeros_like(p)
        nrm[:, :, 2] = 1.0
        _check_product_layout(OwnedDesc(p, nrm, [0] * n, [material(0)], cam))


def test_product_bvh_single_leaf_and_degenerate():
    from tests.helpers import OwnedDesc, material, quad
    cam = ca.camera_look_at((0, 1, 3), (0, 0, 0), (0, 1, 0), 40, 8, 8)
    p, n = quad([-1, 0, -1], [-1, 0, 1], [1, 0, 1], [1, 0, -1])
    od = OwnedDesc(p, n, [0, 1], [material(0), material(5, ka=(1, 1, 1))], cam)
    s = ca.Scene(od.desc)
    info = s.bvh_info()
    assert info.num_nodes == 1 and info.num_leaves == 1
    # 40 copies of the same triangle: centroids coincide, splits fall back to index halves
    P = np.repeat(p[:1], 40, axis=0)
    N = np.repeat(n[:1], 40, axis=0)
    od2 = OwnedDesc(P, N, [0] * 39 + [1], [material(0), material(5, ka=(1, 1, 1))], cam)
    s2 = ca.Scene(od2.desc)
    assert s2.bvh_info().max_leaf_size <= 4


def _check_oracle_tree(mesh):
    a = mesh.arrays()
    s = O.OracleScene(mesh.desc())
    b, m, order = s.export_bvh()
    n = len(a.positions)
    assert sorted(order.tolist()) == list(range(n))
    lo, hi = _tri_bounds(a.positions[order])
    covered = np.zeros(n, np.int32)

    def walk(i, plo, phi):
        nlo, nhi = b[i, :3], b[i, 3:]
        assert (nlo >= plo).all() and (nhi <= phi).all()
        off, cnt, axis = m[i]
        if cnt > 0:
            assert (lo[off:off + cnt] >= nlo).all() and (hi[off:off + cnt] <= nhi).all()
            if cnt > 3:  # only when every centroid in the node coincides
                c = (lo[off:off + cnt] + hi[off:off + cnt]) * 0.5
                assert (np.ptp(c, axis=0) == 0).any()
            covered[off:off + cnt] += 1
            return
        walk(i + 1, nlo, nhi)
        walk(off, nlo, nhi)

    walk(0, np.full(3, -np.inf), np.full(3, np.inf))
    assert (covered == 1).all()


def test_oracle_bvh_invariants():
    _check_oracle_tree(W.cornell().mesh)
    _check_oracle_tree(ca.Mesh.generate("mixed", "tris=10000", 3))


def _random_rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    c = (lo + hi) / 2
    ext = (hi - lo).max()
    o = c + (rng.random((n, 3)) - 0.5) * ext * 1.5
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


@pytest.mark.parametrize("which", ["cornell", "mixed10k"])
def test_oracle_bvh_matches_brute_force(which):
    mesh = W.cornell().mesh if which == "cornell" else ca.Mesh.generate("mixed", "tris=10000", 3)
    a = mesh.arrays()
    lo, hi = a.positions.reshape(-1, 3).min(0), a.positions.reshape(-1, 3).max(0)
    o, d = _random_rays(10000, lo, hi, 5)
    # axis-aligned and grazing rays along the walls of the box
    g = np.random.default_rng(6)
    o2 = (lo + g.random((2000, 3)) * (hi - lo)).astype(np.float32)
    d2 = np.zeros((2000, 3), np.float32)
    d2[np.arange(2000), g.integers(0, 3, 2000)] = g.choice([-1.0, 1.0], 2000)
    d2 += (g.normal(size=(2000, 3)) * 1e-4).astype(np.float32)
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    # rays aimed at random points of random triangles
    k = g.integers(0, len(a.positions), 4000)
    bary = g.dirichlet([1, 1, 1], 4000)
    tgt = np.einsum("nk,nkj->nj", bary, a.positions[k])
    o3 = (tgt + g.normal(size=(4000, 3)) * (hi - lo).max() * 0.3).astype(np.float32)
    d3 = tgt - o3
    d3 = (d3 / np.linalg.norm(d3, axis=1, keepdims=True)).astype(np.float32)
    # exactly axis-aligned rays (two zero direction components: 1/0 in the slab test)
    o4 = (lo + g.random((1500, 3)) * (hi - lo)).astype(np.float32)
    d4 = np.zeros((1500, 3), np.float32)
    d4[np.arange(1500), g.integers(0, 3, 1500)] = g.choice([-1.0, 1.0], 1500)
    o = np.concatenate([o, o2, o3, o4])
    d = np.concatenate([d, d2.astype(np.float32), d3, d4])
    s = O.OracleScene(mesh.desc())
    t1, tt1, u1, v1 = s.trace(o, d)
    t2, tt2, u2, v2 = s.trace(o, d, brute=True)
    assert (t1 >= 0).mean() > 0.3
    assert np.array_equal(t1, t2)
    assert np.array_equal(tt1.view(np.uint32), tt2.view(np.uint32))
    assert np.array_equal(u1, u2) and np.array_equal(v1, v2)


def test_oracle_first_hit_matches_numpy_moller_trumbore():
    """Closest hits from an independent vectorised fp32 Moller-Trumbore (same
    operation order; numpy rounds each operation) over all triangles."""
    wl = W.cornell(24, 16)
    a = wl.mesh.arrays()
    s = O.OracleScene(wl.mesh.desc())
    tri, mat, t = s.first_hit(wl.camera, 24, 16, 2, O.cfg())
    f = np.float32
    v1 = a.positions[:, 0]
    e1 = a.positions[:, 1] - v1
    e2 = a.positions[:, 2] - v1

    def cross(x, y):
        return np.stack([x[..., 1] * y[..., 2] - x[..., 2] * y[..., 1], x[..., 2] * y[..., 0] - x[..., 0] * y[..., 2],
                         x[..., 0] * y[..., 1] - x[..., 1] * y[..., 0]], -1)

    def dot(x, y):
        return (x[..., 0] * y[..., 0] + x[..., 1] * y[..., 1]) + x[..., 2] * y[..., 2]

    assert (tri >= 0).mean() > 0.5 and np.array_equal(tri >= 0, mat >= 0)
    rng = np.random.default_rng(0)
    o = np.repeat(np.array([wl.camera.eye], f), 64, 0)
    dirs = rng.normal(size=(64, 3)).astype(f)
    dirs[:, 0] = -np.abs(dirs[:, 0]) - 1.0
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True).astype(f)
    dirs = dirs.astype(f)
    ids, ts, _, _ = s.trace(o, dirs)
    for i in range(64):
        d = dirs[i][None, :]
        P = cross(d, e2)
        det = dot(e1, P)
        ok = np.abs(det) >= f(1e-5)
        inv = (f(1) / np.where(ok, det, f(1))).astype(f)
        T = (o[i][None, :] - v1).astype(f)
        u = (dot(T, P) * inv).astype(f)
        Q = cross(T, e1)
        v = (dot(d, Q) * inv).astype(f)
        tt = (dot(e2, Q) * inv).astype(f)
        ok &= (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (tt > f(1e-5))
        if not ok.any():
            assert ids[i] == -1
            continue
        best = np.min(tt[ok])
        cand = np.nonzero(ok & (tt == best))[0]
        assert ids[i] == cand.min() and ts[i] == best


def test_stacked_quads_walk_deeper_than_the_lds_stack():
    """The scene of the wavefront's stack-spill test: 40,000 parallel quads, where a
    ray along -z meets every child box at every level. A nearest-first walk must
    hold more than the queue kernels' 16 LDS stack entries when it first reaches a
    leaf, both from the camera and from just above the top quad (shadow rays to
    the light below the stack), so those renders go through the global spill area."""
    from tests.helpers import OwnedDesc, stack_at_first_leaf, stacked_quads
    p, n, m, mats = stacked_quads(40000)
    cam = ca.camera_look_at((0, 0, 2), (0, 0, 0), (0, 1, 0), 40, 64, 48)
    od = OwnedDesc(p, n, m, mats, cam)  # keeps the arrays alive while the scene is built
    s = ca.Scene(od.desc)
    nodes, _, _ = s.export_bvh()
    grid = [(x, y) for x in (-0.5, 0.0, 0.3) for y in (-0.4, 0.1, 0.6)]
    down = [stack_at_first_leaf(nodes, (x, y, 2.0), (0.0, 0.0, -1.0)) for x, y in grid]
    shadow = [stack_at_first_leaf(nodes, (x, y, 1e-5), (-x / 401, -y / 401, -1.0)) for x, y in grid]
    assert min(down) > 16 and min(shadow) > 16, (down, shadow)
