"""The block-Jacobi / RAS partition (precond.cu:414-605) on the host, pinned by
the properties the reference's construction fixes: every RWG in exactly one
block; before bisection a block is exactly the RWGs of one spatial cell of
side max(bbox / (2.5 / radius), mean extent) anchored at the centres'
bounding-box minimum; bisection keeps blocks inside their cell and at most
max_block_rwg large; the overlap of a block is exactly the other RWGs whose
centre lies in the block's bounding box grown by 2 mean extents per layer.
Centres (mean of the 2 Nq quadrature points, bem_fmm.cu:250-300) are
computed here with numpy from the oracle's RWG basis and quadrature rule."""
import numpy as np
import pytest

import bem_amd


@pytest.fixture(scope="module")
def geom(oracle):
    out = {}
    for shape, ref, param in [("sphere", 2, None), ("hex", 1, 0.7)]:
        verts, tris = oracle.make_mesh(shape, ref, param)
        p = oracle.Problem(verts, tris, 7)
        r = p.rwg()
        pts, _ = oracle.quad_rule(7)
        l1, l2 = pts[:, 0], pts[:, 1]
        l0 = 1.0 - l1 - l2

        def half(t):
            v = verts[tris[t]]  # (N, 3 vertices, 3)
            return (l0[None, :, None] * v[:, None, 0, :] + l1[None, :, None] * v[:, None, 1, :] +
                    l2[None, :, None] * v[:, None, 2, :])

        P = np.concatenate([half(r["tri_p"]), half(r["tri_m"])], axis=1)
        c = P.mean(axis=1)
        ext = (P.max(axis=1) - P.min(axis=1)).max(axis=1).mean()
        out[shape] = (verts, tris, c, ext)
        p.close()
    return out


def _cells(c, ext, radius):
    lo = c.min(axis=0)
    cell = max((c.max(axis=0) - lo).max() / (2.5 / radius), ext)
    q = (c - lo) / cell
    # centres within rounding of an interior cell boundary may fall either side
    ambiguous = ((np.abs(q - np.round(q)) < 1e-9) & (np.round(q) >= 1)).any(axis=1)
    return np.floor(q).astype(int), ambiguous


@pytest.mark.parametrize("shape", ["sphere", "hex"])
@pytest.mark.parametrize("radius", [2.0, 4.0])
def test_blocks_are_spatial_cells(geom, shape, radius):
    verts, tris, c, ext = geom[shape]
    parts = bem_amd.block_partition(verts, tris, 7, radius, 10 ** 6, 0)
    N = len(c)
    own = np.concatenate([o for o, _ in parts])
    assert np.array_equal(np.sort(own), np.arange(N))
    cell, amb = _cells(c, ext, radius)
    keys = []
    for o, e in parts:
        assert np.array_equal(o, e) and np.all(np.diff(o) > 0)
        sure = o[~amb[o]]
        assert len(sure) > 0
        k = {tuple(cell[m]) for m in sure}
        assert len(k) == 1, (shape, k)
        keys.append(k.pop())
    assert len(set(keys)) == len(keys)  # one block per cell


@pytest.mark.parametrize("shape,max_block", [("sphere", 40), ("hex", 25), ("sphere", 7)])
def test_bisection_keeps_cells_and_bounds_sizes(geom, shape, max_block):
    verts, tris, c, ext = geom[shape]
    whole = bem_amd.block_partition(verts, tris, 7, 2.0, 10 ** 6, 0)
    parts = bem_amd.block_partition(verts, tris, 7, 2.0, max_block, 0)
    owner = np.zeros(len(c), int)
    for b, (o, _) in enumerate(whole):
        owner[o] = b
    assert np.array_equal(np.sort(np.concatenate([o for o, _ in parts])), np.arange(len(c)))
    for o, _ in parts:
        assert len(o) <= max_block
        assert len(set(owner[o])) == 1  # a bisection never leaves its cell
    assert len(parts) >= int(np.ceil(len(c) / max_block))
    # the halves of a bisected block lie on either side of the midpoint of its longest side
    for o, _ in whole:
        if len(o) <= max_block:
            continue
        lo, hi = c[o].min(axis=0), c[o].max(axis=0)
        ax = int(np.argmax(hi - lo))
        mid = 0.5 * (lo[ax] + hi[ax])
        for p, _ in parts:
            if set(p) <= set(o):
                side = c[p, ax] < mid
                assert side.all() or (~side).all()


@pytest.mark.parametrize("shape", ["sphere", "hex"])
@pytest.mark.parametrize("layers", [1, 2])
def test_overlap_is_the_grown_bounding_box(geom, shape, layers):
    verts, tris, c, ext = geom[shape]
    parts = bem_amd.block_partition(verts, tris, 7, 2.0, 60, layers)
    grow = layers * ext * 2.0
    for o, e in parts:
        assert np.array_equal(e[:len(o)], o)
        ov = e[len(o):]
        assert np.all(np.diff(ov) > 0) and not set(ov) & set(o)
        lo, hi = c[o].min(axis=0) - grow, c[o].max(axis=0) + grow
        inside = np.all((c >= lo) & (c <= hi), axis=1)
        near = np.any(np.minimum(np.abs(c - lo), np.abs(c - hi)) < 1e-12, axis=1)
        want = set(np.nonzero(inside & ~near)[0]) - set(o)
        got = set(ov)
        assert want <= got and got - want <= set(np.nonzero(near)[0])


def test_one_rwg_blocks(geom):
    verts, tris, c, _ = geom["sphere"]
    parts = bem_amd.block_partition(verts, tris, 7, 2.0, 1, 1)
    assert len(parts) == len(c)
    for m, (o, e) in enumerate(parts):
        assert list(o) == [m] and list(e) == [m]


def test_auto_parameters():
    assert bem_amd.auto_precond_params(480, 1.0)["max_block_rwg"] == 1
    p = bem_amd.auto_precond_params(9216, 10.0)
    assert p["overlap_layers"] == 1 and p["max_block_rwg"] == int(np.sqrt(3e8 / 9216))
    assert bem_amd.auto_precond_params(100000, 5.0)["max_block_rwg"] == 100
    assert bem_amd.auto_precond_params(20000, 16.0)["max_block_rwg"] <= 200
    assert bem_amd.auto_precond_params(600, 1.0)["max_block_rwg"] == 600
