"""The checkpoint container read without the library (CPU): a file packed here by hand, with
the layout the library writes, reads back with Julia-order dimensions as [z][y][x] arrays."""

import struct

import numpy as np

import mhdflows_amd as M


def _pack(datasets):
    ents, blobs, off = [], [], 16 + 104 * len(datasets)
    off = (off + 63) // 64 * 64
    for name, code, dims, arr in datasets:
        raw = np.ascontiguousarray(arr).tobytes()
        d4 = list(dims) + [0] * (4 - len(dims))
        ents.append(struct.pack("<48sII4QQQ", name.encode(), code, len(dims), *d4, off, len(raw)))
        blobs.append((off, raw))
        off = (off + len(raw) + 63) // 64 * 64
    buf = bytearray(off)
    buf[0:16] = b"MHDFLOWS" + struct.pack("<II", 1, len(datasets))
    buf[16:16 + 104 * len(ents)] = b"".join(ents)
    for o, raw in blobs:
        buf[o:o + len(raw)] = raw
    return bytes(buf)


def test_read_hand_packed_file(tmp_path):
    nx, ny, nz = 6, 4, 2
    rng = np.random.default_rng(0)
    fields = [rng.standard_normal((nz, ny, nx)) for _ in range(6)]   # [z][y][x] = Julia (x, y, z)
    names = ["i_velocity", "j_velocity", "k_velocity", "i_mag_field", "j_mag_field", "k_mag_field"]
    ds = [(n, 2, (nx, ny, nz), f) for n, f in zip(names, fields)] + [("time", 2, (), np.array([0.125]))]
    sol = (rng.standard_normal((6, nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((6, nz, ny, nx // 2 + 1)))
    ds.append(("sol", 4, (nx // 2 + 1, ny, nz, 6), sol))
    path = tmp_path / "x_t_0003.mhdf"
    path.write_bytes(_pack(ds))
    d = M.read_checkpoint(str(path))
    assert d["time"] == 0.125 and np.array_equal(d["sol"], sol)
    out = M.readMHDFlows(str(path))
    assert len(out) == 7 and out[6] == 0.125
    for a, f in zip(out[:6], fields):
        assert a.dtype == np.float32 and a.shape == (nz, ny, nx)
        assert np.array_equal(a, f.astype(np.float32))
