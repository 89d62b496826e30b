"""Pins of the oracle's cell hashing, sort, reorder and contact search
against brute force (src/sorting.cuh:7-121, src/contactsearch.cuh:12-355)."""

import itertools

import numpy as np
import pytest

from oracle_lib import Oracle, F_CONTACT_OVERFLOW
from sphere_amd.spherebin import Spherebin

pytestmark = pytest.mark.usefixtures("oracle_built")


def _random_state(n=600, L=(0.3, 0.25, 0.28), num=(9, 7, 8), periodic=1, seed=0, r=(0.008, 0.016)):
    rng = np.random.default_rng(seed)
    sb = Spherebin(n, nw=0, sid=f"cells{seed}")
    sb.set_walls(0)
    sb.L = np.asarray(L, float)
    sb.num = np.asarray(num, dtype=np.uint32)
    sb.periodic = periodic
    sb.x = rng.uniform(0, 1, size=(n, 3)) * sb.L
    sb.radius = rng.uniform(*r, size=n)
    sb.k_n = sb.k_t = 1e5
    sb.contactmodel = 2
    sb.time_dt = 1e-6
    sb.time_total = 1.0
    return sb


def _brute_keys(x, L, num, periodic):
    x = x.copy()
    per = [periodic in (1, 2), periodic == 1, False]
    for a in range(3):
        if per[a]:
            x[:, a] = np.where(x[:, a] < 0.0, x[:, a] + L[a], x[:, a])
            x[:, a] = np.where(x[:, a] > L[a], x[:, a] - L[a], x[:, a])
    c = np.floor(x / (L / num.astype(float))).astype(np.int64)
    c = np.clip(c, 0, num.astype(np.int64) - 1)
    return (c[:, 2] * num[1]) * num[0] + c[:, 1] * num[0] + c[:, 0], x


@pytest.mark.parametrize("periodic", [0, 1, 2])
def test_cellid_sort_reorder_vs_brute_force(tmp_path, periodic):
    sb = _random_state(periodic=periodic, seed=periodic)
    # a few positions just outside periodic faces exercise the wrap
    sb.x[:5, 0] = sb.L[0] * 0.999
    o = Oracle.from_spherebin(sb, tmp_path)
    x_in = o.get("x")
    if periodic:
        x_in[0, 0] = sb.L[0] + 1e-4
        x_in[1, 0] = -1e-4
        if periodic == 1:
            x_in[2, 1] = sb.L[1] + 2e-4
        o.set("x", x_in)
    for st in ("cellid", "sort", "reorder"):
        assert o.stage(st) == 0
    keys, xw = _brute_keys(x_in[:, :3], sb.L, sb.num, periodic)
    assert np.array_equal(o.get("cell_keys"), keys.astype(np.uint32))
    assert np.array_equal(o.get("x")[:, :3], xw)
    idx = o.get("sort_idx")
    assert np.array_equal(idx, np.argsort(keys, kind="stable").astype(np.uint32))
    ks = keys[idx]
    K = int(np.prod(sb.num.astype(np.int64)))
    start = np.full(K, 0xFFFFFFFF, dtype=np.uint32)
    end = np.full(K, 0xFFFFFFFF, dtype=np.uint32)
    u, first, counts = np.unique(ks, return_index=True, return_counts=True)
    start[u] = first
    end[u] = first + counts
    assert np.array_equal(o.get("cell_start"), start)
    assert np.array_equal(o.get("cell_end"), end)
    assert np.array_equal(o.get("x_sorted"), o.get("x")[idx])


def _brute_pairs(x, r, L, periodic):
    per = np.array([periodic in (1, 2), periodic == 1, False])
    d = x[:, None, :] - x[None, :, :]
    for a in range(3):
        if per[a]:
            d[..., a] -= L[a] * np.round(d[..., a] / L[a])
    dist = np.sqrt((d ** 2).sum(-1))
    ov = dist < (r[:, None] + r[None, :])
    np.fill_diagonal(ov, False)
    return {(int(i), int(j)) for i, j in zip(*np.nonzero(ov))}


@pytest.mark.parametrize("periodic", [0, 1, 2])
def test_topology_finds_exactly_the_overlaps(tmp_path, periodic):
    """With cells >= one max diameter the slot lists hold exactly the pairs an
    O(N^2) search with periodic images finds."""
    sb = _random_state(n=700, periodic=periodic, seed=10 + periodic, r=(0.012, 0.0165))
    o = Oracle.from_spherebin(sb, tmp_path)
    for st in ("cellid", "sort", "reorder", "topology"):
        assert o.stage(st) == 0
    c = o.get("contacts")
    got = {(i, int(b)) for i in range(sb.np) for b in c[i] if b != sb.np}
    want = _brute_pairs(sb.x, sb.radius, sb.L, periodic)
    assert len(want) > 200
    assert got == want
    # contacts fill from the last slot down: no holes at the top end
    for i in range(sb.np):
        occ = c[i] != sb.np
        k = int(occ.sum())
        assert occ[32 - k:].all()


def test_distmod_is_the_periodic_image(tmp_path):
    sb = _random_state(n=500, periodic=1, seed=4, r=(0.012, 0.0165))
    o = Oracle.from_spherebin(sb, tmp_path)
    for st in ("cellid", "sort", "reorder", "topology"):
        o.stage(st)
    c, dm, x = o.get("contacts"), o.get("distmod"), o.get("x")
    n = sb.np
    checked = 0
    for i in range(n):
        for k in range(32):
            b = c[i, k]
            if b == n:
                continue
            xab = x[i, :3] - x[b, :3] + dm[i, k, :3]
            assert np.linalg.norm(xab) < x[i, 3] + x[b, 3]
            assert dm[i, k, 3] == x[b, 3]
            assert set(np.unique(dm[i, k, :2])) <= {0.0, sb.L[0], -sb.L[0], sb.L[1], -sb.L[1]}
            checked += 1
    assert checked > 100


def _pair_chain(tmp_path, nb=4):
    """Particle 0 overlapping particles 1..nb, all inside one cell block."""
    n = nb + 1
    sb = Spherebin(n, nw=0, sid="slots")
    sb.set_walls(0)
    sb.L = np.array([1.0, 1.0, 1.0])
    sb.num = np.array([10, 10, 10], dtype=np.uint32)
    sb.periodic = 0
    sb.radius[:] = 0.03
    sb.x[0] = (0.5, 0.5, 0.5)
    for j in range(1, n):
        ang = 2 * np.pi * j / nb
        sb.x[j] = (0.5 + 0.05 * np.cos(ang), 0.5 + 0.05 * np.sin(ang), 0.5)
    sb.k_n = sb.k_t = 1e5
    sb.time_dt = 1e-6
    sb.time_total = 1.0
    return Oracle.from_spherebin(sb, tmp_path), sb


def test_slot_rule_keeps_existing_slot_and_fills_highest_hole(tmp_path):
    """src/contactsearch.cuh:252-271: a partner already in a slot keeps it; a
    new partner takes the highest empty slot."""
    o, sb = _pair_chain(tmp_path, nb=4)
    n = sb.np
    c = o.get("contacts")
    c[0, :] = n
    c[0, 3] = 2          # partner 2 already registered low
    c[0, 20] = 4         # partner 4 registered in the middle
    c[0, 31] = 3         # partner 3 in the top slot
    o.set("contacts", c)
    for st in ("cellid", "sort", "reorder", "topology"):
        assert o.stage(st) == 0
    row = o.get("contacts")[0]
    assert row[3] == 2 and row[20] == 4 and row[31] == 3
    # partner 1 took the highest empty slot (30)
    assert row[30] == 1
    assert set(row[row != n]) == {1, 2, 3, 4}


def test_slot_rule_first_match_wins(tmp_path):
    o, sb = _pair_chain(tmp_path, nb=2)
    n = sb.np
    c = o.get("contacts")
    c[0, :] = n
    c[0, 5] = 1
    c[0, 9] = 1          # duplicate entry: the lower one is updated
    o.set("contacts", c)
    d = o.get("distmod")
    d[0, 5] = d[0, 9] = (9.0, 9.0, 9.0, 9.0)
    o.set("distmod", d)
    for st in ("cellid", "sort", "reorder", "topology"):
        o.stage(st)
    d = o.get("distmod")
    assert d[0, 5, 0] == 0.0 and d[0, 9, 0] == 9.0


def test_slot_table_full_is_overflow(tmp_path):
    """Every slot holds another partner and the new overlap finds no slot of
    its own: the reference would write it into slot nc (here 0, the first
    overlap of the search), overwriting a stored contact; this build stops
    with CONTACT_OVERFLOW instead."""
    o, sb = _pair_chain(tmp_path, nb=2)
    c = o.get("contacts")
    c[0, :] = 2          # partner 2 fills every slot: partner 1 finds no room
    o.set("contacts", c)
    for st in ("cellid", "sort", "reorder"):
        o.stage(st)
    assert o.stage("topology") & F_CONTACT_OVERFLOW


def test_more_than_nc_overlaps_is_overflow(tmp_path):
    """A large sphere among > 32 small ones (the reference writes the 33rd
    partner into slot nc = 32, one past its slot array, then stops,
    src/contactsearch.cuh:252-284)."""
    n = 41
    sb = Spherebin(n, nw=0, sid="crowd")
    sb.set_walls(0)
    sb.L = np.array([1.0, 1.0, 1.0])
    sb.num = np.array([4, 4, 4], dtype=np.uint32)
    sb.radius[:] = 0.02
    sb.radius[0] = 0.1
    sb.x[0] = (0.5, 0.5, 0.5)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((n - 1, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    sb.x[1:] = 0.5 + v * 0.115
    sb.k_n = sb.k_t = 1e5
    sb.time_dt = 1e-6
    sb.time_total = 1.0
    o = Oracle.from_spherebin(sb, tmp_path)
    for st in ("cellid", "sort", "reorder"):
        o.stage(st)
    assert o.stage("topology") & F_CONTACT_OVERFLOW
