"""Exact parameter selection (host C++), pinned by the reference's choices.

params.json holds x -> (y, z, k, x_star) and x -> (y, z, c) printed by the
reference build for 375 values of x (powers of ten, 10^n +- 1, 2^63, 2^64,
k-threshold values and 300 seeded log-uniform x).  D(x) depends on these
values, so they must match bit for bit.
"""
import pytest

from tests.conftest import golden


def test_gourdon_params_match_reference(pc):
    rows = golden("params.json")["gourdon"]
    assert len(rows) > 300
    for r in rows:
        got = pc.gourdon_params(int(r["x"]))
        assert got == {k: int(r[k]) for k in ("y", "z", "k", "x_star")}, r["x"]


def test_dr_params_match_reference(pc):
    for r in golden("params.json")["deleglise_rivat"]:
        got = pc.dr_params(int(r["x"]))
        assert got == {k: int(r[k]) for k in ("y", "z", "c")}, r["x"]


def test_params_of_config_sizes(pc):
    for r in golden("configs.json")["D"]:
        got = pc.gourdon_params(int(r["x"]))
        assert got == {k: int(r[k]) for k in ("y", "z", "k", "x_star")}


def test_alpha_overrides(pc):
    try:
        pc.set_alpha_y(2.0)
        pc.set_alpha_z(3.0)
        p = pc.gourdon_params(10 ** 12)
        assert p["y"] == 20000 and p["z"] == 60000
    finally:
        pc.set_alpha_y(0)
        pc.set_alpha_z(0)
    assert pc.gourdon_params(10 ** 12) == {"y": 76500, "z": 76500, "k": 8, "x_star": 1000}


def test_too_large_x_is_rejected(pc):
    with pytest.raises(pc.PrimecountError):
        pc.gourdon_params(1 << 126)
