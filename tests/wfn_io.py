"""Reader of the reference's binary .wfn files (stream access, little-endian;
layout of src/bwfnlib.f90:373-541 header and :124-125 vectors).  Test helper."""
import struct

import numpy as np


def read_wfn(path):
    b = open(path, "rb").read()
    off = struct.unpack_from("<i", b, 12)[0]
    dim = struct.unpack_from("<q", b, off - 12)[0]
    nkeep = struct.unpack_from("<i", b, off - 4)[0]
    o = off
    states = []
    for _ in range(nkeep):
        magic, idx = struct.unpack_from("<2i", b, o)
        e, xj, xt2 = struct.unpack_from("<3f", b, o + 8)
        o += 20
        v = np.frombuffer(b, dtype="<f4", count=dim, offset=o).copy()
        o += 4 * dim
        states.append(dict(magic=magic, index=idx, e=e, xj=xj, xt2=xt2, vec=v))
    return dict(header=b[:off - 4], dim=dim, nkeep=nkeep, states=states, size=len(b), end=o)
