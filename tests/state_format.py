"""The byte form of a state as include/sdx_fem.h documents it, written and
read in plain Python: the tests use it to check the library's encoder and
decoder against the documented layout (the container of the reference's
capture::CaptureState, mochi_physics/src/mochi_capture.cpp:38-75)."""
from __future__ import annotations

import struct

import numpy as np

HEADER = struct.Struct("<4sHH8xQ")   # "mcap", context blocks, entity blocks, 8 zero bytes, bytes after the header
BLOCK = struct.Struct("<QQII")       # type id, entity hash, data size, entity count
FOOTER = b"mend"
HIST_PLANES = 8
GATE = "sdx::CWarmStartGate"
COLLIDER = "sdx::CColliderTarget"


def fnv1a(data: bytes) -> int:
    h = 1469598103934665603
    for b in data:
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def type_id(name: str) -> int:
    return fnv1a(name.encode())


def actor_names(real):
    r = "double" if np.dtype(real) == np.float64 else "float"
    return (f"mochi::CDisplacementSlice<{r}>", f"mochi::CVelocitySlice<{r}>", f"sdx::CNewtonStepHistory<{r}>")


def write(u, v, hist, colliders, gate, real=np.float32, actor_hash=0, collider_hash=0) -> bytes:
    """u, v: one array of 3 nV reals per actor; hist: one array of 8 * 3 nV per actor;
    colliders: one 6-vector per collider; gate: 12 ints."""
    body = BLOCK.pack(type_id(GATE), 0, 48, 0) + np.asarray(gate, np.int32).tobytes()
    n_entity = 0
    if u:
        for name, arrs in zip(actor_names(real), (u, v, hist)):
            data = b"".join(struct.pack("<Q", len(a)) + np.asarray(a, real).tobytes() for a in arrs)
            body += BLOCK.pack(type_id(name), actor_hash, len(data), len(arrs)) + data
            n_entity += 1
    if colliders:
        data = np.asarray(colliders, np.float64).tobytes()
        body += BLOCK.pack(type_id(COLLIDER), collider_hash, len(data), len(colliders)) + data
        n_entity += 1
    body += FOOTER
    return HEADER.pack(b"mcap", 1, n_entity, len(body)) + body


def read(data: bytes) -> dict:
    """Every block of a byte form: {name or hex id: (entity hash, [per-entity arrays or bytes])}."""
    label, n_ctx, n_ent, size = HEADER.unpack_from(data, 0)
    assert label == b"mcap" and size == len(data) - HEADER.size
    names = {type_id(n): n for r in (np.float32, np.float64) for n in actor_names(r)}
    names[type_id(GATE)] = GATE
    names[type_id(COLLIDER)] = COLLIDER
    at = HEADER.size
    out = {}
    for _ in range(n_ctx + n_ent):
        tid, ehash, dsize, count = BLOCK.unpack_from(data, at)
        at += BLOCK.size
        name = names.get(tid, hex(tid))
        chunk = data[at:at + dsize]
        at += dsize
        if name == GATE:
            out[name] = (ehash, np.frombuffer(chunk, np.int32))
        elif name == COLLIDER:
            out[name] = (ehash, np.frombuffer(chunk, np.float64).reshape(count, 6))
        else:
            real = np.float64 if "<double>" in name else np.float32
            arrs, q = [], 0
            for _ in range(count):
                (n,) = struct.unpack_from("<Q", chunk, q)
                q += 8
                arrs.append(np.frombuffer(chunk[q:q + n * np.dtype(real).itemsize], real))
                q += n * np.dtype(real).itemsize
            assert q == dsize
            out[name] = (ehash, arrs)
    assert data[at:at + 4] == FOOTER and at + 4 == len(data)
    return out
