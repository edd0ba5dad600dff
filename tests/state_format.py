"""The reference's binary state file layout, written and read with struct.

save_state_BIN / read_state_BIN (src/stateSR.cpp:54-134): f32 physDim[3],
f32 edge[3], i32 ext[3], f32 l, f32 a, i32 ttlCNT, i32 numCNT, f32 TOL,
i16 occ[G], i32 addr[G] (inclusive prefix of cell capacities), then nine
field-major f32 arrays of ttlCNT slots: x y z cx cy cz l a k.
"""
import struct

import numpy as np

HEADER = struct.Struct("<3f3f3iffiif")


def write(path, box, ext, l, a, cells, tol=1e-6, cap=None):
    """cells: dict cell index -> list of 8-float rods; cap: slots per cell."""
    G = ext[0] * ext[1] * ext[2]
    occ = np.zeros(G, np.int16)
    caps = np.zeros(G, np.int64)
    for c in range(G):
        occ[c] = len(cells.get(c, []))
        caps[c] = max(occ[c], cap[c] if cap is not None else occ[c])
    addr = np.cumsum(caps).astype(np.int32)
    slots = int(addr[-1]) if G else 0
    data = np.zeros((9, slots), np.float32)
    n = 0
    for c in range(G):
        s0 = int(addr[c - 1]) if c else 0
        for j, r in enumerate(cells.get(c, [])):
            data[:8, s0 + j] = r
            data[8, s0 + j] = c
            n += 1
        for j in range(int(occ[c]), int(caps[c])):
            data[8, s0 + j] = -1.0  # empty slot: l = 0
    edge = [box[i] / ext[i] for i in range(3)]
    with open(path, "wb") as f:
        f.write(HEADER.pack(*box, *edge, *ext, l, a, slots, n, tol))
        f.write(occ.tobytes())
        f.write(addr.tobytes())
        f.write(data.tobytes())


def read(path):
    b = open(path, "rb").read()
    h = HEADER.unpack_from(b, 0)
    box, edge, ext = h[0:3], h[3:6], h[6:9]
    l, a, slots, n, tol = h[9:14]
    G = ext[0] * ext[1] * ext[2]
    o = HEADER.size
    occ = np.frombuffer(b, np.int16, G, o)
    addr = np.frombuffer(b, np.int32, G, o + 2 * G)
    data = np.frombuffer(b, np.float32, 9 * slots, o + 6 * G).reshape(9, slots)
    assert len(b) == o + 6 * G + 36 * slots
    return {"box": box, "edge": edge, "ext": ext, "len": l, "radius": a, "slots": slots,
            "n_rods": n, "tol": tol, "occ": occ, "addr": addr, "data": data}
