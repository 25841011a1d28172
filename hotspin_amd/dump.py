"""The `#dump002` binary frame format of hotspin field output.

Layout (src/hotspin-core/dump/writer.go:46-97, reader.go:31-133): eight 8-byte
words -- magic "#dump002", components, mesh size (3 words, internal ZYX
order), mesh step (3 float64, internal order), mesh unit (string), time
(float64), time unit, data label, data unit (8-byte NUL-padded strings),
precision (4 = float32) -- then components*N float32 values in internal
order (component 0 = user z), then a CRC-64 (ISO polynomial, Go's
hash/crc64) of everything before it, 0 meaning "not set".
"""
import struct

import numpy as np

MAGIC = b"#dump002"

_ISO_POLY = 0xD800000000000000


def _make_table():
    t = np.zeros(256, dtype=np.uint64)
    for i in range(256):
        crc = i
        for _ in range(8):
            crc = (crc >> 1) ^ _ISO_POLY if crc & 1 else crc >> 1
        t[i] = crc
    return [int(x) for x in t]


_TABLE = _make_table()


def crc64_iso(data, crc=0):
    """CRC-64/ISO as computed by Go's hash/crc64 with crc64.ISO."""
    tab = _TABLE
    crc = ~crc & 0xFFFFFFFFFFFFFFFF
    for b in bytes(data):
        crc = tab[(crc ^ b) & 0xFF] ^ (crc >> 8)
    return ~crc & 0xFFFFFFFFFFFFFFFF


def _str8(s):
    b = s.encode() if isinstance(s, str) else bytes(s)
    return (b + b"\0" * 8)[:8]


class Frame:
    def __init__(self, data, mesh_step=(1.0, 1.0, 1.0), mesh_unit="m", time=0.0,
                 time_unit="s", label="", unit=""):
        self.data = np.asarray(data, dtype=np.float32)
        assert self.data.ndim == 4, "data must be (components, N0, N1, N2)"
        self.mesh_step = tuple(float(x) for x in mesh_step)
        self.mesh_unit = mesh_unit
        self.time = float(time)
        self.time_unit = time_unit
        self.label = label
        self.unit = unit
        self.crc = None

    @property
    def components(self):
        return self.data.shape[0]

    @property
    def mesh_size(self):
        return tuple(self.data.shape[1:])


def read(path, check_crc=True):
    with open(path, "rb") as f:
        buf = f.read()
    return loads(buf, check_crc=check_crc)


def loads(buf, check_crc=True):
    off = 0

    def word():
        nonlocal off
        w = buf[off:off + 8]
        off += 8
        return w

    magic = word().rstrip(b"\0")
    if magic != MAGIC:
        raise ValueError("dump: bad magic number: %r" % magic)
    comp = struct.unpack("<Q", word())[0]
    size = [struct.unpack("<Q", word())[0] for _ in range(3)]
    step = [struct.unpack("<d", word())[0] for _ in range(3)]
    mesh_unit = word().split(b"\0")[0].decode()
    time = struct.unpack("<d", word())[0]
    time_unit = word().split(b"\0")[0].decode()
    label = word().split(b"\0")[0].decode()
    unit = word().split(b"\0")[0].decode()
    prec = struct.unpack("<Q", word())[0]
    if prec != 4:
        raise ValueError("dump: unsupported precision %d" % prec)
    n = comp * size[0] * size[1] * size[2]
    data = np.frombuffer(buf, dtype="<f4", count=n, offset=off).reshape(comp, *size)
    off += 4 * n
    crc = struct.unpack("<Q", buf[off:off + 8])[0]
    if check_crc and crc != 0:
        mine = crc64_iso(buf[:off])
        if mine != crc:
            raise ValueError("dump CRC error: expected %016x, got %016x" % (crc, mine))
    fr = Frame(data.copy(), step, mesh_unit, time, time_unit, label, unit)
    fr.crc = crc
    return fr


def dumps(frame, crc=True):
    d = frame.data
    out = bytearray()
    out += _str8(MAGIC)
    out += struct.pack("<Q", d.shape[0])
    for s in d.shape[1:]:
        out += struct.pack("<Q", s)
    for s in frame.mesh_step:
        out += struct.pack("<d", s)
    out += _str8(frame.mesh_unit)
    out += struct.pack("<d", frame.time)
    out += _str8(frame.time_unit)
    out += _str8(frame.label)
    out += _str8(frame.unit)
    out += struct.pack("<Q", 4)
    out += np.ascontiguousarray(d, dtype="<f4").tobytes()
    out += struct.pack("<Q", crc64_iso(out) if crc else 0)
    return bytes(out)


def write(path, frame, crc=True):
    with open(path, "wb") as f:
        f.write(dumps(frame, crc))
