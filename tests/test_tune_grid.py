"""The tuner's record of the tile grid (msa_amd/tune.py): every point of
{32, 64, 128} rows per lane x {16, 32, 64} lanes per pair is either realised
by a measured kernel shape or marked dominated with its reason and evidence."""
from msa_amd import tune


def fake_measured():
    def rec(ms):
        return {"ms": ms, "gcups": 1000.0 / ms, "equal_to_auto": True, "routing": {}}
    shapes = {"auto": rec(50.0), "batched_16x32": rec(50.0), "batched_16x16": rec(51.0),
              "batched_32x16": rec(55.0), "long": dict(rec(90.0), rows_per_lane=8, profile_slots=21),
              "plain": rec(180.0)}
    return {"uneven-2000": shapes, "dna-64x30000": dict(shapes, long=dict(rec(110.0), rows_per_lane=32))}


RES = {"batched_16x32": {"vgprs": 168}, "batched_16x16": {"vgprs": 109}, "batched_32x16": {"vgprs": 132},
       "long_8": {"vgprs": 58}, "long_16": {"vgprs": 76}, "long_32": {"vgprs": 109}}


def test_every_grid_point_is_measured_or_dominated():
    g = tune.grid_map(fake_measured(), RES)
    pts = {(p["rows_per_lane"], p["lanes_per_pair"]): p for p in g["points"]}
    assert set(pts) == {(w, l) for w in (32, 64, 128) for l in (16, 32, 64)}
    assert pts[(32, 16)]["status"] == "measured" and pts[(32, 16)]["shape"] == "batched_16x32"
    assert pts[(32, 16)]["timings"]["uneven-2000"]["ms"] == 50.0
    assert pts[(32, 64)]["shape"] == "long"
    for key, p in pts.items():
        if key in ((32, 16), (32, 64)):
            continue
        assert p["status"] == "dominated" and p["reason"] and p["evidence"]
    # 32 lanes x 32 rows: a 1 KiB profile block per code, and the measured
    # cost of 32-lane groups.
    p = pts[(32, 32)]
    assert "1024 bytes per code" in p["reason"]
    assert p["evidence"]["ms_32x16_over_16x16"]["uneven-2000"] == round(55.0 / 51.0, 4)
    # 128 rows per lane: registers beyond what a wave addresses.
    assert pts[(128, 16)]["evidence"]["estimated_vgprs"] > 256
    assert "beyond the 256" in pts[(128, 64)]["reason"]


def test_grid_without_resource_counts():
    g = tune.grid_map(fake_measured(), {})
    assert all(p["status"] for p in g["points"])
    assert all(p.get("evidence", {}).get("estimated_vgprs") is None for p in g["points"] if p["status"] == "dominated")
