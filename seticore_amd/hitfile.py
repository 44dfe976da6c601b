"""Reader for .hits and .events files: streams of unpacked Cap'n Proto messages
with the schema of hit.capnp (Hit{signal: Signal, filterbank: Filterbank},
Event{hits: List(Hit)}).

Plays the role of the reference's hitls tool (hitls.cpp:19-58). Pure Python,
no capnp dependency. Struct sizes are read from the pointers, so files written
with the older schema (fewer fields, e.g. data/voyager.hits) decode too:
fields beyond a struct's data section read as 0, as Cap'n Proto specifies.
"""

import struct

SIGNAL_FIELDS = [  # (name, byte offset, format) - hit.capnp:11-46
    ("frequency", 0, "<d"), ("index", 8, "<i"), ("drift_steps", 12, "<i"),
    ("drift_rate", 16, "<d"), ("snr", 24, "<f"), ("coarse_channel", 28, "<i"),
    ("beam", 32, "<i"), ("num_timesteps", 36, "<i"), ("power", 40, "<f"),
    ("incoherent_power", 44, "<f"),
]
FILTERBANK_FIELDS = [  # hit.capnp:50-78 (data fields; pointers: source_name, data)
    ("fch1", 0, "<d"), ("foff", 8, "<d"), ("tstart", 16, "<d"), ("tsamp", 24, "<d"),
    ("ra", 32, "<d"), ("dec", 40, "<d"), ("telescope_id", 48, "<i"),
    ("num_timesteps", 52, "<i"), ("num_channels", 56, "<i"), ("coarse_channel", 60, "<i"),
    ("start_channel", 64, "<i"), ("beam", 68, "<i"),
]


class _Msg:
    def __init__(self, segments):
        self.segs = segments

    def word(self, seg, i):
        return struct.unpack_from("<Q", self.segs[seg], 8 * i)[0]

    def resolve(self, seg, i):
        """Follows the pointer at (seg, word i); returns (kind, seg, target, ptr)."""
        p = self.word(seg, i)
        kind = p & 3
        if kind == 2:  # far pointer
            landing = (p & 0xFFFFFFFF) >> 3
            seg2 = p >> 32
            if (p >> 2) & 1:
                raise ValueError("double-far pointers are not produced by this schema")
            return self.resolve(seg2, landing)
        off = (p & 0xFFFFFFFF) >> 2
        if off & (1 << 29):
            off -= 1 << 30
        return kind, seg, i + 1 + off, p

    def struct_at(self, seg, i):
        if self.word(seg, i) == 0:
            return None
        kind, s, t, p = self.resolve(seg, i)
        assert kind == 0
        dw = (p >> 32) & 0xFFFF
        pw = (p >> 48) & 0xFFFF
        return s, t, dw, pw

    def list_at(self, seg, i):
        if self.word(seg, i) == 0:
            return None
        kind, s, t, p = self.resolve(seg, i)
        assert kind == 1
        esize = (p >> 32) & 7
        count = p >> 35
        return s, t, esize, count


def _fields(msg, st, fields):
    s, t, dw, _ = st
    out = {}
    buf = msg.segs[s]
    for name, off, fmt in fields:
        size = struct.calcsize(fmt)
        out[name] = struct.unpack_from(fmt, buf, 8 * t + off)[0] if off + size <= 8 * dw else 0
    return out


def read_messages(data):
    """Yields each message as a list of segment byte strings."""
    pos = 0
    while pos < len(data):
        nseg = struct.unpack_from("<I", data, pos)[0] + 1
        sizes = struct.unpack_from("<%dI" % nseg, data, pos + 4)
        hdr = 4 + 4 * nseg
        hdr += (8 - hdr % 8) % 8
        pos += hdr
        segs = []
        for sz in sizes:
            segs.append(data[pos:pos + 8 * sz])
            pos += 8 * sz
        yield segs


def decode_hit(segments):
    """Decodes one Hit message into {'signal': {...}, 'filterbank': {...}}."""
    msg = _Msg(segments)
    out = _decode_hit_struct(msg, msg.struct_at(0, 0))
    out["layout"] = {"segments": [len(x) // 8 for x in segments]}
    return out


def _decode_hit_struct(msg, root):
    import numpy as np
    s, t, dw, pw = root
    out = {"signal": None, "filterbank": None}
    sig = msg.struct_at(s, t + dw + 0) if pw > 0 else None
    if sig is not None:
        out["signal"] = _fields(msg, sig, SIGNAL_FIELDS)
    fb = msg.struct_at(s, t + dw + 1) if pw > 1 else None
    if fb is not None:
        f = _fields(msg, fb, FILTERBANK_FIELDS)
        fs, ft, fdw, fpw = fb
        name = msg.list_at(fs, ft + fdw + 0) if fpw > 0 else None
        if name is not None:
            ls, lt, _, cnt = name
            raw = msg.segs[ls][8 * lt:8 * lt + cnt]
            f["source_name"] = raw.rstrip(b"\0").decode("utf-8", "replace")
        else:
            f["source_name"] = ""
        dl = msg.list_at(fs, ft + fdw + 1) if fpw > 1 else None
        if dl is not None:
            ls, lt, esize, cnt = dl
            assert esize == 4
            f["data"] = np.frombuffer(msg.segs[ls], dtype="<f4", count=cnt, offset=8 * lt).copy()
        else:
            f["data"] = np.zeros(0, np.float32)
        out["filterbank"] = f
    return out


def decode_event(segments):
    """Decodes one Event message (hit.capnp:91-93) into a list of hits, one per
    cadence file ({'signal': {...} or None, 'filterbank': {...}})."""
    msg = _Msg(segments)
    s, t, dw, pw = msg.struct_at(0, 0)
    ls, lt, esize, words = msg.list_at(s, t + dw)
    assert esize == 7  # inline composite
    tag = msg.word(ls, lt)
    count = (tag & 0xFFFFFFFF) >> 2
    edw = (tag >> 32) & 0xFFFF
    epw = (tag >> 48) & 0xFFFF
    assert count * (edw + epw) == words
    return [_decode_hit_struct(msg, (ls, lt + 1 + i * (edw + epw), edw, epw))
            for i in range(count)]


def read_events(path_or_bytes):
    """Decodes every event in a .events file (path) or byte string."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        data = bytes(path_or_bytes)
    else:
        with open(path_or_bytes, "rb") as f:
            data = f.read()
    return [decode_event(segs) for segs in read_messages(data)]


def read_hits(path_or_bytes):
    """Decodes every hit in a .hits file (path) or byte string."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        data = bytes(path_or_bytes)
    else:
        with open(path_or_bytes, "rb") as f:
            data = f.read()
    return [decode_hit(segs) for segs in read_messages(data)]
