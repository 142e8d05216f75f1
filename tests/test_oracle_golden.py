"""The CPU oracle against the reference's own outputs (golden fixtures from
oracle/_ref/crystal2_ref, tests/golden/make_golden.py) and closed forms."""
import os

import numpy as np
import pytest

import crystal2_amd as c2
import pyoracle
from conftest import GOLDEN, relmse

SCN = os.path.join(GOLDEN, "scenes")


def load(name):
    d = os.path.join(SCN, name)
    return c2.Scene.load(os.path.join(d, "scene.json"), d)


@pytest.mark.parametrize("name,spp,depth", [("cornell", 4, 5), ("zoo_nophong", 8, 6), ("zoo_extra", 8, 6),
                                            ("disks", 8, 6), ("textured", 8, 6)])
def test_oracle_film_bit_identical_to_reference(name, spp, depth):
    """libm variant + clang argument order = the reference binary, bit for bit."""
    sc = load(name)
    film, _ = pyoracle.OracleScene(sc, variant="libm").render(64, 64, spp, depth, arg_order="ltr")
    ref = np.load(os.path.join(GOLDEN, f"film_{name}.npy"))
    assert np.array_equal(film, ref)


def test_oracle_textured_phong_statistical():
    """Blinn-Phong with a Kd texture: statistical agreement (the redraw generator), and
    every pixel the Phong surfaces do not reach stays bit-identical."""
    d = os.path.join(SCN, "textured")
    sc = c2.Scene.load(os.path.join(d, "scene_phong.json"), d)
    film, _ = pyoracle.OracleScene(sc, variant="libm").render(64, 64, 8, 6, arg_order="ltr")
    ref = np.load(os.path.join(GOLDEN, "film_textured_phong.npy"))
    assert relmse(film, ref) < 2e-3
    assert (film == ref).all(-1).mean() > 0.8


def test_oracle_phong_statistical():
    """Blinn-Phong's rejection redraws use the reference's shared mt19937; the oracle's
    deterministic substitute agrees statistically only."""
    sc = load("zoo")
    film, _ = pyoracle.OracleScene(sc, variant="libm").render(64, 64, 8, 6, arg_order="ltr")
    ref = np.load(os.path.join(GOLDEN, "film_zoo.npy"))
    assert relmse(film, ref) < 2e-3
    assert (film == ref).all(-1).mean() > 0.9


def test_portable_math_variant_close_to_reference():
    """Transcendentals evaluated in double and rounded: ulp-level differences only."""
    sc = load("zoo_nophong")
    film, _ = pyoracle.OracleScene(sc, variant="portable").render(64, 64, 8, 6, arg_order="ltr")
    ref = np.load(os.path.join(GOLDEN, "film_zoo_nophong.npy"))
    assert relmse(film, ref) < 1e-4


@pytest.mark.parametrize("name", ["icosphere", "zoo", "disks"])
def test_oracle_hits_match_reference(name):
    g = np.load(os.path.join(GOLDEN, f"hits_{name}.npz"))
    sc = load(name)
    o = pyoracle.OracleScene(sc, accel=pyoracle.ACCEL_BVH, tie=pyoracle.TIE_FIRST, variant="libm")
    ids, t = o.intersect(g["rays"])
    assert np.array_equal(ids, g["ids"])
    hit = ids >= 0
    assert np.array_equal(t[hit], g["t"][hit])
    anyh = o.intersect_test(g["rays"])
    assert np.array_equal(anyh, g["any"].astype(bool))


def test_brute_matches_bvh_except_ties():
    """Brute force (first entity in scene order wins) vs the median BVH."""
    g = np.load(os.path.join(GOLDEN, "hits_icosphere.npz"))
    sc = load("icosphere")
    b = pyoracle.OracleScene(sc, accel=pyoracle.ACCEL_BRUTE, tie=pyoracle.TIE_LOWEST)
    ids, t = b.intersect(g["rays"])
    hit = g["ids"] >= 0
    assert np.array_equal(t[hit], g["t"][hit])
    assert (ids == g["ids"]).mean() >= 0.9999 - 1e-12 or np.array_equal(t[ids != g["ids"]], g["t"][ids != g["ids"]])


@pytest.mark.parametrize("wh", [(256, 256), (1920, 1080), (64, 64)])
def test_halton_matches_reference(wh):
    W, H = wh
    g = np.load(os.path.join(GOLDEN, f"halton_{W}x{H}.npz"))
    idx, vals = pyoracle.halton(W, H, g["queries"], 100)
    assert np.array_equal(idx, g["index"])
    assert np.array_equal(vals, g["values"])
    # the product's host implementation (used to build the device tables)
    for i in range(0, len(g["queries"]), 37):
        x, y, k = (int(v) for v in g["queries"][i])
        assert c2.halton_index(W, H, x, y, k) == g["index"][i]
        for d in (0, 1, 2, 7, 50, 99):
            assert c2.halton_value(W, H, int(g["index"][i]), d) == g["values"][i, d]


def test_halton_closed_forms():
    """Constants and radical inverses that arithmetic fixes (HaltonSampler.cpp:204-229)."""
    # (128, 243) scales, (7, 5) exponents, stride 31104, inverses 243^-1 mod 128 = 59, 128^-1 mod 243 = 131
    assert (243 * 59) % 128 == 1 and (128 * 131) % 243 == 1
    W = H = 256
    # dim 0 is the base-2 radical inverse of idx >> 7; dim 1 the base-3 one of idx / 243
    for idx in (0, 1, 5, 128, 129, 31104, 5845119):
        v0 = c2.halton_value(W, H, idx, 0)
        b = idx >> 7
        exact0 = sum(((b >> i) & 1) * 2.0 ** -(i + 1) for i in range(32))
        assert v0 == np.float32(exact0)
        v = idx // 243
        exact1, inv, digits = 0.0, 1.0 / 3, 0
        while v:
            exact1 += (v % 3) * inv
            inv /= 3
            v //= 3
            digits += 1
        # fp32 evaluation: one rounding of invBase per digit plus the final product
        v1 = c2.halton_value(W, H, idx, 1)
        assert 0 <= v1 < 1 and abs(v1 - exact1) <= (digits + 2) * 2.0 ** -24
    # pixel offsets are below stride * 188 (no "% stride" in the reference)
    assert max(c2.halton_index(W, H, x, y, 0) for x in range(0, 128, 7) for y in range(0, 128, 5)) < 5845120


def test_permutation_tables_are_std_shuffle():
    """Dims 2..99 use std::shuffle(iota(p), mt19937()) tables: value of index 1 in dim d is
    (perm[1] + perm[0] * invB / (1 - invB)) * invB for the table the reference built."""
    g = np.load(os.path.join(GOLDEN, "halton_256x256.npz"))
    assert np.all(g["values"] < 1) and np.all(g["values"] >= 0)
    # first table (base 5) is a permutation: the values at indices 1..4 are distinct
    vals = [c2.halton_value(256, 256, i, 2) for i in range(1, 5)]
    assert len(set(vals)) == 4


def test_pow5_product_equals_rounded_double_pow():
    """Schlick's x^5 is evaluated as three double products rounded once to float
    (host and device alike); on 2^20 arguments in [0, 1] it equals pow in double
    rounded to float."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.random(1 << 20).astype(np.float32),
                        np.array([0, 1, 0.5, 1e-8, 0.999999], np.float32)])
    d = x.astype(np.float64)
    d2 = d * d
    prod = (d2 * d2 * d).astype(np.float32)
    ref = np.power(d, 5.0).astype(np.float32)
    assert np.array_equal(prod, ref)


@pytest.mark.parametrize("name,exact", [("cornell", True), ("zoo_nophong", True), ("zoo", False),
                                        ("zoo_extra", True), ("disks", True), ("textured", True)])
def test_oracle_direct_lighting_matches_reference(name, exact):
    """DirectLightingIntegrator films of the reference binary (film_*_dl.npy): bit-identical
    for the libm variant; with Phong the rejection redraws agree statistically only."""
    sc = load(name)
    spp = 4 if name == "cornell" else 8
    film, _ = pyoracle.OracleScene(sc, variant="libm", integrator="direct").render(64, 64, spp, 6, arg_order="ltr")
    ref = np.load(os.path.join(GOLDEN, f"film_{name}_dl.npy"))
    if exact:
        assert np.array_equal(film, ref)
    else:
        assert relmse(film, ref) < 1e-6
        assert (film == ref).all(-1).mean() > 0.99


@pytest.mark.parametrize("name", ["zoo", "disks"])
def test_oracle_kdtree_hits_match_reference(name):
    """"Acceleration": "KDTree": the reference's KD-tree hits, which differ from its BVH's
    (the KD-tree keeps a farther hit in ~0.3% of these rays)."""
    g = np.load(os.path.join(GOLDEN, f"hits_{name}_kd.npz"))
    o = pyoracle.OracleScene(load(name), accel=pyoracle.ACCEL_KDTREE, variant="libm")
    ids, t = o.intersect(g["rays"])
    assert np.array_equal(ids, g["ids"])
    assert np.array_equal(t[ids >= 0], g["t"][ids >= 0])
    assert np.array_equal(o.intersect_test(g["rays"]), g["any"].astype(bool))
    bvh = np.load(os.path.join(GOLDEN, f"hits_{name}.npz"))
    assert (g["ids"] != bvh["ids"]).any()  # the fixture tells the two accelerators apart


@pytest.mark.parametrize("name,spp,depth,integ,exact", [
    ("cornell", 4, 5, "path", True), ("zoo_nophong", 8, 6, "path", True), ("disks", 8, 6, "path", True),
    ("textured", 8, 6, "path", True), ("cornell", 4, 5, "direct", True), ("disks", 8, 6, "direct", True),
    ("textured", 8, 6, "direct", True), ("zoo_nophong", 8, 6, "direct", False)])
def test_oracle_kdtree_films_match_reference(name, spp, depth, integ, exact):
    """KD-tree films of the reference binary. Direct lighting on the zoo reads an
    uninitialised pdf in ~1% of pixels there (see the direct-lighting test): statistical."""
    o = pyoracle.OracleScene(load(name), accel=pyoracle.ACCEL_KDTREE, variant="libm", integrator=integ)
    film, _ = o.render(64, 64, spp, depth, arg_order="ltr")
    ref = np.load(os.path.join(GOLDEN, f"film_{name}{'_dl' if integ == 'direct' else ''}_kd.npy"))
    if exact:
        assert np.array_equal(film, ref)
    else:
        assert (film == ref).all(-1).mean() > 0.98
        assert relmse(film, ref) < 1e-3
