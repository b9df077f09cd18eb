"""The byte and JSON forms of a state, on the CPU (no world, no device):
the library's decoder, IsEqualState and JSON conversion against byte forms
written in plain Python from the layout include/sdx_fem.h documents (the
container of the reference's capture::CaptureState,
mochi_physics/src/mochi_capture.cpp:38-75, 456-826).

- two soft actors and two colliders written by tests/state_format.py convert
  to JSON with every value read back exactly, names "Soft", "Soft2",
  "Rigid", "Rigid2" as the reference's CaptureStateToJson names actors of
  one type;
- IsEqualState: equal values are equal whatever their bytes (0 and -0, NaN
  and NaN), a changed value, gate counter, collider target or entity hash is
  not;
- malformed bytes (label, footer, sizes, unknown component, disagreeing
  actor blocks) are rejected with SDX_ERR_INVALID_ARGUMENT.
"""
import json
import struct

import numpy as np
import pytest

from project_superdex_amd._lib import SdxError
from project_superdex_amd.scene import is_equal_state_bytes, state_bytes_to_json

import state_format as sf


def sample(real=np.float32, seed=0):
    rng = np.random.default_rng(seed)
    nv = (5, 3)
    u = [rng.standard_normal(3 * n).astype(real) for n in nv]
    v = [rng.standard_normal(3 * n).astype(real) for n in nv]
    hist = [rng.standard_normal(sf.HIST_PLANES * 3 * n).astype(real) for n in nv]
    colliders = [rng.standard_normal(6) for _ in range(2)]
    gate = list(range(1, 13))
    return u, v, hist, colliders, gate


@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_json_of_a_documented_byte_form(real):
    u, v, hist, colliders, gate = sample(real)
    b = sf.write(u, v, hist, colliders, gate, real, actor_hash=7, collider_hash=9)
    doc = json.loads(state_bytes_to_json(b))
    assert list(doc["actors"]) == ["Soft", "Soft2", "Rigid", "Rigid2"]
    disp, vel, hst = sf.actor_names(real)
    for a, name in enumerate(["Soft", "Soft2"]):
        act = doc["actors"][name]
        assert act["_handle"] == a and act["_type"] == "Soft" and act["_name"] == ""
        c = act["components"]
        np.testing.assert_array_equal(np.array(c[disp], real), u[a])
        np.testing.assert_array_equal(np.array(c[vel], real), v[a])
        np.testing.assert_array_equal(np.array(c[hst], real).ravel(), hist[a])
    for k, name in enumerate(["Rigid", "Rigid2"]):
        t = doc["actors"][name]["components"][sf.COLLIDER]
        np.testing.assert_array_equal(t["target"] + t["stage_start_target"], colliders[k])
    g = doc["scene"][sf.GATE]
    assert g["skip"] + g["seen"] + g["net_skip"] == gate
    # the indented form holds the same document
    assert json.loads(state_bytes_to_json(b, pretty=True)) == doc
    assert "\n" in state_bytes_to_json(b, pretty=True) and "\n" not in state_bytes_to_json(b)


def test_json_of_a_world_without_actors_and_non_finite_values():
    b = sf.write([], [], [], [], [0] * 12)
    assert json.loads(state_bytes_to_json(b)) == {"scene": {sf.GATE: {"skip": [0] * 4, "seen": [0] * 4,
                                                                   "net_skip": [0] * 4}}}
    u, v, hist, colliders, gate = sample()
    u[0][1] = np.nan
    v[1][0] = np.inf
    doc = json.loads(state_bytes_to_json(sf.write(u, v, hist, colliders, gate)))
    disp, vel, _ = sf.actor_names(np.float32)
    assert doc["actors"]["Soft"]["components"][disp][1] is None
    assert doc["actors"]["Soft2"]["components"][vel][0] is None


def test_is_equal_state_compares_values_not_bytes():
    u, v, hist, colliders, gate = sample()
    a = sf.write(u, v, hist, colliders, gate, actor_hash=3)
    assert is_equal_state_bytes(a, a)
    # 0 and -0, NaN and NaN (other payload bits) are equal values
    u0 = [x.copy() for x in u]
    u0[0][0] = 0.0
    u1 = [x.copy() for x in u]
    u1[0][0] = -0.0
    u0[1][2] = np.float32(np.nan)
    u1[1][2] = np.frombuffer(struct.pack("<I", 0x7FC00001), np.float32)[0]
    b0 = sf.write(u0, v, hist, colliders, gate, actor_hash=3)
    b1 = sf.write(u1, v, hist, colliders, gate, actor_hash=3)
    assert b0 != b1 and is_equal_state_bytes(b0, b1)

    def differs(**kw):
        args = dict(u=u, v=v, hist=hist, colliders=colliders, gate=gate, actor_hash=3)
        args.update(kw)
        return not is_equal_state_bytes(a, sf.write(args["u"], args["v"], args["hist"], args["colliders"],
                                                    args["gate"], actor_hash=args["actor_hash"]))

    v2 = [x.copy() for x in v]
    v2[1][4] += 1e-3
    h2 = [x.copy() for x in hist]
    h2[0][-1] *= 2
    c2 = [x.copy() for x in colliders]
    c2[1][5] += 1.0
    assert differs(v=v2) and differs(hist=h2) and differs(colliders=c2)
    assert differs(gate=gate[::-1]) and differs(actor_hash=4)
    assert not is_equal_state_bytes(a, sf.write(u, v, hist, colliders, gate, np.float64, actor_hash=3))


def test_malformed_bytes_are_rejected():
    u, v, hist, colliders, gate = sample()
    good = sf.write(u, v, hist, colliders, gate)
    state_bytes_to_json(good)
    bad = {
        "label": b"mcaq" + good[4:],
        "footer": good[:-4] + b"mEnd",
        "short": good[:-8],
        "long": good + b"\0",
        "size": good[:16] + struct.pack("<Q", len(good)) + good[24:],
        "no bytes": b"",
    }
    # an unknown component type
    unknown = bytearray(good)
    struct.pack_into("<Q", unknown, sf.HEADER.size, 12345)
    bad["unknown type"] = bytes(unknown)
    # actor blocks that disagree on their actors (velocities of other lengths)
    v_short = [x[:-3] for x in v]
    bad["actors disagree"] = sf.write(u, v_short, hist, colliders, gate)
    # an array length that is no multiple of 3 reals per vertex
    bad["length"] = sf.write([x[:-1] for x in u], [x[:-1] for x in v], [x[:-8] for x in hist], colliders, gate)
    for name, b in bad.items():
        with pytest.raises(SdxError) as ei:
            state_bytes_to_json(b)
        assert ei.value.code == -1, name


def test_python_reader_reads_the_python_writer():
    """The restatement's writer and reader agree (so a GPU test can read a world's bytes with it)."""
    u, v, hist, colliders, gate = sample(np.float64)
    blocks = sf.read(sf.write(u, v, hist, colliders, gate, np.float64, actor_hash=5, collider_hash=6))
    disp, vel, hst = sf.actor_names(np.float64)
    assert blocks[disp][0] == 5 and blocks[sf.COLLIDER][0] == 6
    for a in range(2):
        np.testing.assert_array_equal(blocks[disp][1][a], u[a])
        np.testing.assert_array_equal(blocks[vel][1][a], v[a])
        np.testing.assert_array_equal(blocks[hst][1][a], hist[a])
    np.testing.assert_array_equal(blocks[sf.COLLIDER][1], np.array(colliders))
    np.testing.assert_array_equal(blocks[sf.GATE][1], gate)
