"""The disk basis cache in ion_gym's entry format (io/fa_cache.py,
io/basis_cache.py) and the post-solve flight conditions
(physics/flight_conditions.py): key text and hashes recomputed by hand,
round trips, corruption and race handling, the voltage-free geometry key,
and a funnel deck re-tuned through set_flight_conditions onto cached bases,
composed bit for bit like the direct composition."""
import hashlib
import json
import os
import shutil

import numpy as np
import pytest

from ion_gym_amd import configs as C
from ion_gym_amd.io import basis_cache as B
from ion_gym_amd.io import fa_cache as F
from ion_gym_amd.physics import fields_rz as FR
from ion_gym_amd.physics.flight_conditions import (FlightConditionError,
                                                   set_flight_conditions,
                                                   solved_key)
from ion_gym_amd.physics.flight_spec import CollisionSpec


def _sha16(b):
    return hashlib.sha256(b).hexdigest()[:16]


# ---------------------------------------------------------------- fa_cache
def test_key_is_the_hash_of_the_canonical_text():
    assert F.spec_key({"name": "x", "h": 1.0}) == \
        _sha16(b'{"spec":{"h":1.0,"name":"x"},"v":2}')
    nested = {"b": [1, 2.5, {"z": None, "a": True}], "a": "s"}
    assert F.spec_key(nested) == _sha16(
        b'{"spec":{"a":"s","b":[1,2.5,{"a":true,"z":null}]},"v":2}')
    # representation noise below 1e-10 keeps the key, a real change moves it
    assert F.spec_key({"h": 1.0 + 1e-15}) == F.spec_key({"h": 1.0})
    assert F.spec_key({"h": 0.1 + 0.2}) == F.spec_key({"h": 0.3})
    assert F.spec_key({"h": 0.5}) != F.spec_key({"h": 1.0})
    # numpy scalars and arrays key as the Python values
    assert F.spec_key({"v": np.float64(2.0), "n": np.int64(3),
                       "a": np.array([1.0, 2.0])}) == \
        F.spec_key({"v": 2.0, "n": 3, "a": [1.0, 2.0]})


def test_round_trip_manifest_and_read_only_memmaps(tmp_path):
    phi = np.linspace(-3.0, 7.0, 3000).reshape(50, 60)
    lab = (np.arange(3000) % 7).astype(np.uint8).reshape(50, 60)
    spec = {"name": "rt", "h": 0.1}
    key = F.store(spec, {"phi": phi, "lab": lab}, root=tmp_path)
    assert key == F.spec_key(spec)
    assert sorted(p.name for p in tmp_path.iterdir()) == [key]
    meta = json.loads((tmp_path / key / "meta.json").read_text())
    assert meta["key"] == key and meta["version"] == 2 and \
        meta["spec"] == spec
    # sha_head: the first 4096 raw bytes of the array data
    assert meta["arrays"]["phi"] == {
        "shape": [50, 60], "dtype": "float64",
        "sha_head": _sha16(phi.tobytes()[:4096])}
    assert meta["arrays"]["lab"]["sha_head"] == _sha16(lab.tobytes()[:4096])
    arrs, meta2 = F.load(spec, root=tmp_path)
    assert meta2 == meta
    assert np.array_equal(arrs["phi"], phi) and arrs["phi"].dtype == phi.dtype
    assert np.array_equal(arrs["lab"], lab)
    assert isinstance(arrs["phi"], np.memmap)
    assert not arrs["phi"].flags.writeable
    # by key string too, and a subset of names
    arrs, _ = F.load(key, root=tmp_path, names=["lab"])
    assert list(arrs) == ["lab"]
    assert F.verify(spec, root=tmp_path)
    # a second store of the same key keeps the entry
    assert F.store(spec, {"phi": phi * 2}, root=tmp_path) == key
    assert np.array_equal(F.load(key, root=tmp_path)[0]["phi"], phi)


def test_miss_version_and_key_mismatch(tmp_path):
    assert F.load({"name": "none"}, root=tmp_path) == (None, None)
    key = F.store({"name": "v"}, {"a": np.ones(3)}, root=tmp_path)
    mp = tmp_path / key / "meta.json"
    meta = json.loads(mp.read_text())
    meta["version"] = 1
    mp.write_text(json.dumps(meta))
    assert F.load(key, root=tmp_path) == (None, None)
    assert F.verify(key, root=tmp_path) is False


def test_truncated_and_flipped_arrays_are_refused(tmp_path):
    spec = {"name": "trunc"}
    key = F.store(spec, {"Ez": np.arange(5000.0), "phi": np.arange(64.0)},
                  root=tmp_path)
    ez = tmp_path / key / "Ez.npy"
    with open(ez, "r+b") as fh:
        fh.truncate(os.path.getsize(ez) // 2)
    with pytest.raises(IOError, match="truncated or corrupt"):
        F.load(spec, root=tmp_path)
    assert F.verify(spec, root=tmp_path) is False
    # a flipped first data byte: same shape and dtype, wrong head hash
    key = F.store({"name": "hash"}, {"phi": np.arange(64.0)}, root=tmp_path)
    npy = tmp_path / key / "phi.npy"
    raw = bytearray(npy.read_bytes())
    raw[-64 * 8] ^= 0xFF
    npy.write_bytes(bytes(raw))
    with pytest.raises(IOError, match="head hash mismatch"):
        F.load(key, root=tmp_path)
    # an array the manifest lists but the directory lacks
    key = F.store({"name": "gone"}, {"a": np.ones(2), "b": np.ones(2)},
                  root=tmp_path)
    (tmp_path / key / "b.npy").unlink()
    with pytest.raises(IOError, match="missing array b"):
        F.load(key, root=tmp_path)


def test_incomplete_entry_is_replaced(tmp_path, capsys):
    spec = {"name": "partial"}
    key = F.store(spec, {"a": np.ones(4), "b": np.zeros(4)}, root=tmp_path)
    (tmp_path / key / "b.npy").unlink()
    F.store(spec, {"a": np.full(4, 2.0), "b": np.full(4, 3.0)},
            root=tmp_path)
    assert "INCOMPLETE" in capsys.readouterr().out
    arrs, _ = F.load(key, root=tmp_path)
    assert np.array_equal(arrs["b"], np.full(4, 3.0))


def test_a_lost_race_leaves_no_temp_copy(tmp_path, monkeypatch):
    real = os.replace

    def lose(src, dst):
        if os.path.basename(str(src)).startswith(".") and os.path.isdir(src):
            shutil.copytree(src, dst)        # the other writer publishes
            raise OSError(66, "Directory not empty")
        return real(src, dst)

    monkeypatch.setattr(F.os, "replace", lose)
    key = F.store({"name": "race"}, {"phi": np.arange(12.0).reshape(3, 4)},
                  root=tmp_path)
    assert sorted(p.name for p in tmp_path.iterdir()) == [key]
    arrs, _ = F.load(key, root=tmp_path, mmap=False)
    assert np.array_equal(arrs["phi"], np.arange(12.0).reshape(3, 4))


def test_provenance_inventory_and_removal(tmp_path):
    k1 = F.store({"name": "one"}, {"a": np.ones((3, 5))}, root=tmp_path)
    k2 = F.store({"name": "two"}, {"a": np.ones(7), "b": np.ones(2)},
                 root=tmp_path)
    F.annotate(k1, "deckA", label="r-z | 3x5", root=tmp_path,
               spec_json={"x": 1})
    F.annotate(k1, "deckA", label="r-z | 3x5", root=tmp_path)
    F.annotate(k1, "deckB", root=tmp_path)
    F.annotate(k1, "", root=tmp_path)                 # no-op
    F.annotate("0" * 16, "ghost", root=tmp_path)      # absent entry: no-op
    recs = json.loads((tmp_path / k1 / "provenance.json").read_text())
    assert [r["name"] for r in recs] == ["deckA", "deckB"]
    assert recs[0]["spec_json"] == {"x": 1}
    (tmp_path / "solve_rates.json").write_text("{}")  # a loose file
    rows = {r["key"]: r for r in F.entries(root=tmp_path)}
    assert set(rows) == {k1, k2}
    assert rows[k1]["produced_by"] == ["deckA", "deckB"]
    assert rows[k1]["descriptor"] == "r-z | 3x5"
    assert rows[k1]["has_spec_json"] is True
    assert rows[k1]["spec_name"] == "one" and rows[k1]["shape"] == [3, 5]
    assert rows[k2]["arrays"] == 2 and rows[k2]["problems"] == []
    assert rows[k1]["bytes"] == sum(
        os.path.getsize(tmp_path / k1 / f) for f in os.listdir(tmp_path / k1))
    (tmp_path / k2 / "provenance.json").write_text("not json")
    rows = {r["key"]: r for r in F.entries(root=tmp_path)}
    assert rows[k2]["problems"] and "provenance.json" in \
        rows[k2]["problems"][0]
    assert F.remove(k1, root=tmp_path)[0] is True
    assert F.remove(k1, root=tmp_path) == (False, 0)
    n, freed = F.clear_all(root=tmp_path)
    assert n == 1 and freed > 0
    assert sorted(p.name for p in tmp_path.iterdir()) == ["solve_rates.json"]


def test_default_root_follows_the_environment(tmp_path, monkeypatch):
    monkeypatch.setenv("ION_GYM_CACHE", str(tmp_path / "c1"))
    assert F.DEFAULT_ROOT == str(tmp_path / "c1")
    key = F.store({"name": "env"}, {"a": np.ones(2)})
    assert (tmp_path / "c1" / key / "meta.json").exists()
    monkeypatch.setenv("ION_GYM_CACHE", str(tmp_path / "c2"))
    assert F.load(key) == (None, None)


# ---------------------------------------------------------------- basis_cache
def test_rz_bases_round_trip_in_the_reference_layout(tmp_path):
    spec = C.funnel_deck(coarse=8)
    lab = C.rz_deck_labels(spec)
    rng = np.random.default_rng(0)
    bases = {i: rng.normal(size=lab.shape) for i in (1, 2, 3, 100)}
    key = B.store(spec, bases, lab, root=tmp_path)
    assert key == B.key(spec)
    meta = json.loads((tmp_path / key / "meta.json").read_text())
    assert meta["arrays"]["ele"]["dtype"] == "uint8"
    assert {n for n in meta["arrays"]} == {"b1", "b2", "b3", "b100", "ele"}
    assert all(meta["arrays"][f"b{i}"]["dtype"] == "float64"
               for i in bases)
    got, ele = B.load(spec, root=tmp_path)
    assert ele.dtype == np.int16 and np.array_equal(ele, lab)
    assert set(got) == set(bases)
    for i in bases:
        assert np.array_equal(got[i], bases[i])
        assert got[i].flags.writeable
    rows = F.entries(root=tmp_path)
    assert rows[0]["descriptor"] == \
        "2-D/rz | 151x26 @ 0.8 mm/gu | 100 electrodes"
    # the store (with its descriptor) and the hit (without) each record
    assert rows[0]["produced_by"] == ["funnel_deck_c8", "funnel_deck_c8"]
    assert rows[0]["has_spec_json"] is True


def test_geometry_key_is_voltage_free():
    base = C.funnel_deck(coarse=4)
    k0 = B.key(base)
    for mutate in (
            lambda s: setattr(s.geometry.electrodes[3], "dc", 42.0),
            lambda s: setattr(s.geometry.electrodes[3], "name", "renamed"),
            lambda s: setattr(s.geometry.electrodes[3], "color", [1, 2, 3]),
            lambda s: setattr(s.geometry.electrodes[3], "rf_groups", []),
            lambda s: setattr(s.geometry.electrodes[3], "dc_index", 77),
            lambda s: setattr(s.geometry.electrodes[3], "dc_group", None),
            lambda s: setattr(s.geometry.dc_groups[0], "v_in", 300.0),
            lambda s: setattr(s.geometry.rf_groups[0], "amplitude_v", 1.0),
            lambda s: setattr(s.integration, "dt_ns", 0.5)):
        s = C.funnel_deck(coarse=4)
        mutate(s)
        assert B.key(s) == k0
    for mutate in (
            lambda s: setattr(s.geometry, "mm_per_gu", 0.2),
            lambda s: setattr(s.geometry, "symmetry", "planar"),
            lambda s: s.geometry.electrodes[5].shapes[0].update(r0_mm=2.0),
            lambda s: setattr(s.geometry.electrodes[5], "basis", 500),
            lambda s: s.geometry.electrodes.reverse()):
        s = C.funnel_deck(coarse=4)
        mutate(s)
        assert B.key(s) != k0
    # the solver code is part of the key
    kd = B.cache_key_dict(base)
    assert kd["_solve_code"] == B.solve_code_key()
    assert B.key(base) == F.spec_key(kd)


def test_boolean_labels_are_a_loud_miss_and_wide_labels_refused(tmp_path,
                                                                 capsys):
    spec = C.funnel_deck(coarse=8)
    lab = C.rz_deck_labels(spec)
    F.store(B.cache_key_dict(spec), {"b1": np.zeros(lab.shape),
                                     "ele": lab > 0}, root=tmp_path)
    assert B.load(spec, root=tmp_path) == (None, None)
    assert "pre-format-2" in capsys.readouterr().out
    with pytest.raises(ValueError, match="8-bit label array"):
        B.store(spec, {1: np.zeros(lab.shape)}, np.full(lab.shape, 256),
                root=tmp_path / "w")


# ---------------------------------------------------------------- conditions
@pytest.fixture(scope="module")
def funnel4():
    """The coarse-4 funnel deck, its labels and bases solved once."""
    from ion_gym_amd.physics.solver2d import solve_rz
    spec = C.funnel_deck(coarse=4)
    lab = C.rz_deck_labels(spec)
    keys = sorted(int(v) for v in np.unique(lab) if v)
    bases = dict(zip(keys, solve_rz(lab, [{k: 1e4} for k in keys],
                                    spec.geometry.mm_per_gu)))
    return spec, lab, bases


def test_deck_labels_and_volts_match_the_funnel_config(funnel4):
    spec, lab, bases = funnel4
    lab_c, h, bases_c, dc_c, members_c, _groups = C.funnel_electrodes(4)
    assert np.array_equal(lab, lab_c) and h == spec.geometry.mm_per_gu
    dc, members, groups = FR.spec_terms_rz(spec)
    assert dc == dc_c and members == members_c
    assert [g.name for g in groups] == ["rfA", "rfB"]
    for k in bases:
        assert np.array_equal(bases[k], bases_c[k])


def test_retune_reuses_cached_bases_and_composes_exactly(funnel4, tmp_path,
                                                          capsys):
    spec, lab, bases = funnel4
    spec = C.funnel_deck(coarse=4)
    spec.collisions = CollisionSpec()
    B.store(spec, bases, lab, root=tmp_path)
    sk = solved_key(spec)
    lines = set_flight_conditions(
        spec, dc_groups={"ladder": {"v_in": 150.0, "v_out": 3.0}},
        drives={"rfB": {"amplitude_v": 80.0, "phase_deg": 90.0}},
        voltages={}, ion={"ke_ev": (0.5, 1.5), "mz_list": [400.0]},
        gas={"p_torr": 3.0, "T_k": 310.0},
        integration={"dt_ns": 0.5}, solved_key=sk)
    out = capsys.readouterr().out
    assert "FLIGHT CONDITIONS" in out and "ladder.v_in: 100.0 -> 150.0" in out
    assert len(lines) == 9
    assert spec.collisions.P_torr == 3.0 and spec.collisions.T_k == 310.0
    assert spec.source.ke_lo == 0.5 and spec.integration.dt_ns == 0.5
    got, ele = B.load(spec, root=tmp_path)
    assert got is not None, "a re-tune must hit the cached bases"
    model = FR.compose_spec_rz(spec, got, ele.astype(np.int32))
    # the direct composition with the same volts, written out by hand
    dc = {i + 1: 150.0 + (3.0 - 150.0) * i / 99 for i in range(100)}
    members = {i + 1: ("rfA",) if i % 2 == 0 else ("rfB",)
               for i in range(100)}
    groups = [FR.DriveGroup("rfA", "sin", 1.0e6, 100.0, 0.0),
              FR.DriveGroup("rfB", "sin", 1.0e6, 80.0, 90.0)]
    ref = FR.compose_rz(bases, lab, spec.geometry.mm_per_gu, dc, members,
                        groups)
    for name in ("EzA", "EuA", "EzK", "EuK", "ch_om", "ch_ph"):
        assert np.array_equal(getattr(model, name), getattr(ref, name)), name


def test_flight_condition_refusals(funnel4):
    spec = C.funnel_deck(coarse=4)
    spec.collisions = CollisionSpec()
    with pytest.raises(FlightConditionError, match="unknown electrode"):
        set_flight_conditions(spec, voltages={"nope": 1.0}, show=False)
    with pytest.raises(FlightConditionError,
                       match="non-uniform dc ladder 'ladder'"):
        set_flight_conditions(spec, voltages={"ring007": 1.0}, show=False)
    with pytest.raises(FlightConditionError, match="unknown group 'rfC'"):
        set_flight_conditions(spec, drives={"rfC": {"amplitude_v": 1}},
                              show=False)
    with pytest.raises(FlightConditionError, match="has no field 'ampl'"):
        set_flight_conditions(spec, drives={"rfA": {"ampl": 1}}, show=False)
    with pytest.raises(FlightConditionError, match="settable"):
        set_flight_conditions(spec, dc_groups={"ladder": {"vin": 1}},
                              show=False)
    with pytest.raises(FlightConditionError, match="source has no field"):
        set_flight_conditions(spec, ion={"mass": 1}, show=False)
    with pytest.raises(FlightConditionError, match="P_pa is deprecated"):
        set_flight_conditions(spec, gas={"P_pa": 100.0}, show=False)
    with pytest.raises(FlightConditionError, match="pressure once"):
        set_flight_conditions(spec, gas={"p_torr": 1.0, "P_torr": 1.0},
                              show=False)
    with pytest.raises(FlightConditionError, match="integration has no"):
        set_flight_conditions(spec, integration={"steps": 3}, show=False)
    # a uniform group moves all its members
    spec.geometry.dc_groups[0].uniform = True
    lines = set_flight_conditions(spec, voltages={"ring007": 12.0},
                                  show=False)
    assert "also moves 99 other member(s)" in lines[0]
    assert all(e.dc == 12.0 for e in spec.geometry.electrodes)
    # a geometry change made outside the call, caught by the solved key
    sk = solved_key(spec)
    spec.geometry.electrodes[0].shapes[0]["r0_mm"] = 11.0
    with pytest.raises(FlightConditionError,
                       match="does not match the solved fields"):
        set_flight_conditions(spec, integration={"dt_ns": 1.0},
                              solved_key=sk, show=False)
