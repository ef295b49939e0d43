"""Frame packets of the GPU service protocol.

The reference's service sends a JSON `ready` message with its backend and
feature list when a browser connects (gpu/server.py:76), then one binary
packet per frame (gpu/server.py:17-20): u32 little-endian header length, a
compact UTF-8 JSON header, spaces up to a multiple of 4 bytes, and 8
little-endian float32 per previewed particle (gpu/solver.py:986-1032). The
browser worker checks the sizes before it renders
(src/three/gpu-worker.js:36-42). This module builds those messages from the
engine's device snapshot (mpm_snapshot) and the library's packet builder
(mpm_frame_packet), and decodes them with the worker's checks.
"""
from __future__ import annotations

import ctypes as C
import json
import struct

import numpy as np

from . import lib as L

# Features the engine can honour: jittered solid sampling and configurable
# discretisation. Phase change is not part of the engine, so the worker's
# melting scenes are refused on its side, as with an older service.
FEATURES = ("solidSampling", "discretization")


def ready_message(device="MI355X", arch="hip", features=FEATURES):
    """The handshake text frame the worker waits for."""
    return json.dumps(dict(type="ready", backend=dict(arch=arch, device=device), features=list(features)))


def header_bytes(meta):
    """Compact UTF-8 JSON, no NaN or infinity, as the service encodes it."""
    return json.dumps(meta, separators=(",", ":"), ensure_ascii=False, allow_nan=False).encode("utf-8")


def packet(meta, data):
    """One frame packet (bytes) of header `meta` and rows `data` [n, 8]."""
    head = header_bytes(meta)
    rows = np.ascontiguousarray(np.asarray(data, np.float32).reshape(-1, 8))
    lib = L.load()
    length = C.c_uint64()
    fp = rows.ctypes.data_as(C.POINTER(C.c_float))
    L.check(lib.mpm_frame_packet(head, len(head), fp, len(rows), None, 0, C.byref(length)))
    out = (C.c_uint8 * length.value)()
    L.check(lib.mpm_frame_packet(head, len(head), fp, len(rows), out, length.value, C.byref(length)))
    return bytes(out)


def decode(buf):
    """(header dict, rows [n, 8] float32) of a packet, with the worker's
    checks: the header fits, count and renderCount are integers with
    0 <= renderCount <= count <= capacity, and the packet ends exactly after
    renderCount rows. Raises ValueError otherwise."""
    buf = bytes(buf)
    if len(buf) < 4:
        raise ValueError("packet shorter than its length field")
    (n,) = struct.unpack_from("<I", buf)
    if n > len(buf) - 4:
        raise ValueError("header length past the end of the packet")
    meta = json.loads(buf[4:4 + n].decode("utf-8"))
    offset = 4 + (n + 3) // 4 * 4
    count, capacity = meta.get("count"), meta.get("capacity")
    rendered = meta.get("renderCount", count)
    ints = all(isinstance(v, int) and not isinstance(v, bool) for v in (count, rendered))
    if not ints or rendered < 0 or rendered > count or count < 0 or count > capacity or \
            offset + rendered * 32 != len(buf):
        raise ValueError("invalid particle frame")
    return meta, np.frombuffer(buf, "<f4", rendered * 8, offset).reshape(rendered, 8)


def frame(engine, preview_limit=100_000, codes=None, **extra):
    """(header dict, rows) of the engine's current state with the totals the
    service reports (gpu/solver.py:1024-1030); `extra` adds fields such as
    time, substeps or domain. Raises RuntimeError on a non-finite state, as
    the service does."""
    st, data = engine.snapshot(preview_limit, codes)
    if st["nonfinite"] or not np.isfinite(data).all() or not np.isfinite([st["mass"], st["min_J"]]).all():
        raise RuntimeError("Non-finite GPU state; simulation stopped")
    meta = dict(type="frame", count=st["count"], capacity=int(engine.capacity), renderCount=st["render_count"],
                sampleStride=st["sample_stride"], models=st["models"], mass=st["mass"], energy=0.0,
                solids=st["solids"], liquids=st["liquids"], min=st["min_temperature"], max=st["max_temperature"],
                minJ=st["min_J"], maxStress=st["max_stress"], maxPlastic=st["max_plastic"])
    meta.update(extra)
    return meta, data
