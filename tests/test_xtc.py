"""GROMACS XTC frames (the reference's trajectory source through xdrfile,
calcIR.cu:212-219, 341-395): the XDR frame layout, the compressed coordinate
code (round trips at several precisions, the large-range path, water runs),
frame subsets, appends and corrupt input.

No XTC file written by GROMACS is available in this environment, so the
decoder is pinned by (a) a byte-exact hand-assembled uncompressed frame, (b)
compressed frames assembled bit by bit from the published xdr3dfcoord layout
(mixed-radix triples, run flags, runs of small differences with the
first-pair exchange, run lengths kept under a flag bit 0, the small-triple
precision raised and lowered between runs, the per-axis large-range path),
independent of this project's writer, and (c) round trips through the
writer.  Agreement with files that
GROMACS itself wrote stays parity unpinned (none exist here)."""

import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import calcir_amd as C


def _uncompressed_frame(natoms, step, time, box, xyz):
    b = struct.pack(">iiif", 1995, natoms, step, time)
    b += struct.pack(">9f", *box)
    b += struct.pack(">i", natoms)
    b += struct.pack(">%df" % (3 * natoms), *xyz.ravel())
    return b


def test_uncompressed_frame_layout(tmp_path):
    rng = np.random.default_rng(0)
    f = tmp_path / "small.xtc"
    frames = []
    with open(f, "wb") as fh:
        for k in range(3):
            xyz = rng.uniform(-5, 5, (7, 3)).astype(np.float32)
            box = [2.5 + k, 0.1, 0.2, 0.3, 3.5 + k, 0.4, 0.5, 0.6, 4.5 + k]
            fh.write(_uncompressed_frame(7, 100 * k, 0.25 * k, box, xyz))
            frames.append((xyz, box))
    assert C.xtc_scan(f) == (7, 3)
    xyz, box, times, steps = C.xtc_read(f)
    for k in range(3):
        assert np.array_equal(xyz[k], frames[k][0])          # exact: stored as floats
        b = frames[k][1]
        assert np.array_equal(box[k], np.float32([b[0], b[4], b[8]]))  # the diagonal
    assert np.array_equal(times, np.float32([0.0, 0.25, 0.5]))
    assert list(steps) == [0, 100, 200]


@pytest.mark.parametrize("nwater,precision", [(3, 1000.0), (64, 1000.0), (1024, 1000.0), (256, 100.0),
                                              (256, 10000.0)])
def test_round_trip_water(tmp_path, nwater, precision):
    xyz, box, times = C.synth_trajectory(nwater, 4, 11 + nwater)
    f = tmp_path / "w.xtc"
    C.xtc_write(f, xyz, box, times, precision=precision)
    assert C.xtc_scan(f) == (4 * nwater, 4)
    x2, b2, t2, s2 = C.xtc_read(f)
    natoms = 4 * nwater
    if natoms <= 9:
        assert np.array_equal(x2, xyz)
    else:
        # x*precision rounded half away from zero in FP32, then scaled back
        lf = xyz * np.float32(precision)
        q = np.trunc(lf + np.where(lf >= 0, np.float32(0.5), np.float32(-0.5)))
        assert np.array_equal(x2, q.astype(np.float32) * np.float32(1.0 / precision))
        assert np.abs(x2 - xyz).max() <= 0.5 / precision + 1e-6
    assert np.array_equal(b2, box) and np.array_equal(t2, times) and list(s2) == [0, 1, 2, 3]


def test_water_runs_compress(tmp_path):
    # O, H1, H2, M within 0.1 nm of each other: the small-difference runs
    # must carry most atoms (well under 12 bytes per atom)
    xyz, box, times = C.synth_trajectory(1024, 1, 5)
    f = tmp_path / "c.xtc"
    C.xtc_write(f, xyz, box, times)
    size = os.path.getsize(f)
    assert size < 4.5 * 4096, size / 4096


def test_large_range_path_and_scattered_atoms(tmp_path):
    rng = np.random.default_rng(3)
    # a range of 2e4 nm at precision 1000 exceeds 2^24 integers per axis:
    # coordinates are coded per axis (no mixed radix)
    xyz = rng.uniform(-1e4, 1e4, (2, 50, 3)).astype(np.float32)
    box = np.full((2, 3), 1e4, np.float32)
    times = np.float32([0.0, 0.002])
    f = tmp_path / "big.xtc"
    C.xtc_write(f, xyz, box, times)
    x2 = C.xtc_read(f)[0]
    assert np.abs(x2 - xyz).max() <= 0.5 / 1000 + 2e-3  # float spacing at 1e4 nm is ~1e-3
    # scattered atoms (no runs) in a normal box
    xyz = rng.uniform(0, 3, (1, 300, 3)).astype(np.float32)
    C.xtc_write(f, xyz, np.full((1, 3), 3.0, np.float32), np.float32([0.0]))
    assert np.abs(C.xtc_read(f)[0] - xyz).max() <= 0.5 / 1000 + 1e-6


def test_subset_append_and_errors(tmp_path):
    xyz, box, times = C.synth_trajectory(32, 6, 9)
    f = tmp_path / "a.xtc"
    C.xtc_write(f, xyz[:4], box[:4], times[:4])
    C.xtc_write(f, xyz[4:], box[4:], times[4:], append=True)
    assert C.xtc_scan(f) == (128, 6)
    full = C.xtc_read(f)[0]
    part, _, t, _ = C.xtc_read(f, first=2, count=3)
    assert np.array_equal(part, full[2:5]) and np.array_equal(t, times[2:5])
    # truncated file: the frames before the cut are found, the cut one is corrupt
    data = open(f, "rb").read()
    g = tmp_path / "cut.xtc"
    open(g, "wb").write(data[: len(data) - 100])
    with pytest.raises(C.CalcirError):
        C.xtc_scan(g)
    with pytest.raises(C.CalcirError) as e:
        C.xtc_read(f, first=4, count=3)
    assert e.value.status == 6  # a missing frame, as a missing desired time
    h = tmp_path / "junk.xtc"
    open(h, "wb").write(b"not an xtc file at all")
    with pytest.raises(C.CalcirError):
        C.xtc_scan(h)
    with pytest.raises(C.CalcirError) as e:
        C.xtc_scan(tmp_path / "missing.xtc")
    assert e.value.status == 2


@pytest.mark.parametrize("nwater", [2, 64])  # uncompressed (natoms <= 9) and compressed frames
def test_header_pass_and_sequential_blocks(tmp_path, nwater):
    # cir_xtc_times (coordinate blocks skipped by their byte counts) and the
    # sequential reader (cir_xtc_open / _next) give what cir_xtc_read gives
    xyz, box, times = C.synth_trajectory(nwater, 23, 17)
    tri = np.zeros((23, 3, 3), np.float32)
    tri[:, 0, 0], tri[:, 1, 1], tri[:, 2, 2] = box[:, 0], box[:, 1], box[:, 2]
    tri[:, 1, 0] = 0.1 * box[:, 0]  # an off-diagonal element to carry through
    f = tmp_path / "b.xtc"
    C.xtc_write(f, xyz, tri, times)
    xr, br, tr, _ = C.xtc_read(f, full_box=True)
    th, bh = C.xtc_times(f)
    assert np.array_equal(th, tr) and np.array_equal(bh, br)
    got = list(C.xtc_blocks(f, 7))
    assert [g[0].shape[0] for g in got] == [7, 7, 7, 2]
    assert np.array_equal(np.concatenate([g[0] for g in got]), xr)
    assert np.array_equal(np.concatenate([g[1] for g in got]), br)
    assert np.array_equal(np.concatenate([g[2] for g in got]), tr)
    # reading past the end is a missing frame
    L = C.lib()
    import ctypes
    rd = ctypes.c_void_p()
    C.check(L.cir_xtc_open(os.fsencode(f), xyz.shape[1], ctypes.byref(rd)))
    buf = np.zeros((24,) + xyz.shape[1:], np.float32)
    try:
        assert L.cir_xtc_next(rd, 24, buf.ctypes.data, None, None) == 6
    finally:
        L.cir_xtc_close(rd)


def test_frame_offsets_and_seek(tmp_path):
    # random access as the reference's xtc_seek + xdr_seek: the offset index
    # comes from headers and byte counts; a seek then reads what cir_xtc_read
    # reads from that frame
    xyz, box, times = C.synth_trajectory(64, 17, 23)
    f = tmp_path / "s.xtc"
    C.xtc_write(f, xyz, box, times)
    off = C.xtc_offsets(f)
    assert off[0] == 0 and np.all(np.diff(off) > 0) and off[-1] < os.path.getsize(f)
    xr, br, tr, _ = C.xtc_read(f, full_box=True)
    for first in (11, 3, 0, 16):
        got = list(C.xtc_blocks(f, 4, first=first))
        assert np.array_equal(np.concatenate([g[0] for g in got]), xr[first:])
        assert np.array_equal(np.concatenate([g[2] for g in got]), tr[first:])
    # a range read skips the earlier frames without decoding them
    part, _, t, _ = C.xtc_read(f, first=9, count=5)
    assert np.array_equal(part, xr[9:14]) and np.array_equal(t, tr[9:14])
    import ctypes
    L = C.lib()
    rd = ctypes.c_void_p()
    C.check(L.cir_xtc_open(os.fsencode(f), xyz.shape[1], ctypes.byref(rd)))
    try:
        assert L.cir_xtc_seek(rd, 17) == 6
        C.check(L.cir_xtc_seek(rd, 5))  # backwards after forwards
        buf = np.zeros((2,) + xyz.shape[1:], np.float32)
        C.check(L.cir_xtc_next(rd, 2, buf.ctypes.data, None, None))
        assert np.array_equal(buf, xr[5:7])
    finally:
        L.cir_xtc_close(rd)


def _patched(tmp_path, name, data, offset, value):
    b = bytearray(data)
    b[offset:offset + 4] = struct.pack(">i", value)
    g = tmp_path / name
    g.write_bytes(bytes(b))
    return g


@pytest.mark.timeout(60)
def test_corrupt_headers_and_coordinate_blocks_are_errors(tmp_path):
    # A compressed frame: 52 header bytes (magic, natoms, step, time, box),
    # then natoms, precision, minint[3] at 60, maxint[3] at 72, smallidx,
    # and the byte count at 88 (maxint[2] < minint[2] makes the z radix 0,
    # a divisor of the decode).  Corrupt values end as errors: no
    # backward seek (a negative atom count), no zero radix (maxint < minint),
    # no allocation sized by garbage (byte or atom counts beyond the file).
    xyz, box, times = C.synth_trajectory(16, 3, 5)
    f = tmp_path / "ok.xtc"
    C.xtc_write(f, xyz, box, times)
    data = f.read_bytes()
    assert struct.unpack(">i", data[4:8])[0] == 64
    nbytes = struct.unpack(">i", data[88:92])[0]
    frame = 92 + (nbytes + 3) // 4 * 4
    assert len(data) == 3 * frame
    L = C.lib()
    off = np.zeros(3, np.int64)
    t = np.zeros(3, np.float32)
    for name, at, value in (("radix.xtc", 80, struct.unpack(">i", data[68:72])[0] - 1),
                            ("nbytes.xtc", 88, 0x7FFFFF00),
                            ("huge.xtc", 4, 0x7FFFFFFF)):
        g = _patched(tmp_path, name, data, at, value)
        with pytest.raises(C.CalcirError):
            C.xtc_scan(g)
        if name != "huge.xtc":
            rd = np.zeros((1, 64, 3), np.float32)
            assert L.cir_xtc_read(os.fsencode(g), 0, 1, 64, rd.ctypes.data, None, None, None) != 0
        if name == "nbytes.xtc":  # the header-only passes skip by the byte count
            assert L.cir_xtc_offsets(os.fsencode(g), 3, off.ctypes.data) != 0
            assert L.cir_xtc_times(os.fsencode(g), 3, t.ctypes.data, None) != 0
    # a negative atom count in the second frame's header: the passes stop there
    g = _patched(tmp_path, "neg.xtc", data, frame + 4, -5)
    assert C.xtc_scan(g) == (64, 1)
    assert L.cir_xtc_offsets(os.fsencode(g), 1, off.ctypes.data) == 0 and off[0] == 0
    assert L.cir_xtc_offsets(os.fsencode(g), 2, off.ctypes.data) != 0
    assert L.cir_xtc_times(os.fsencode(g), 2, t.ctypes.data, None) != 0
    # the untouched file still reads
    assert C.xtc_scan(f) == (64, 3)


def test_cli_rejects_a_corrupt_trajectory_without_crashing(tmp_path):
    # the drop-in executable reads the trajectory before it touches the GPU:
    # a coordinate block whose z range is empty ends the run with
    # EXIT_FAILURE and a message, not a signal
    import subprocess
    xyz, box, times = C.synth_trajectory(16, 3, 5)
    f = tmp_path / "ok.xtc"
    C.xtc_write(f, xyz, box, times)
    data = f.read_bytes()
    g = _patched(tmp_path, "bad.xtc", data, 80, struct.unpack(">i", data[68:72])[0] - 1)
    inp = tmp_path / "input.inp"
    inp.write_text("xtcf %s\noutf %s\nmodel tip4p2005\nspecies H2O\nntcfpoints 2\nnsamples 1\n" % (g, tmp_path / "o"))
    r = subprocess.run([C.CLI_PATH, str(inp)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=60)
    assert r.returncode == 1, (r.returncode, r.stdout[-500:])
    assert "could not be opened" in r.stdout
    assert not list(tmp_path.glob("o_*.dat"))


@pytest.mark.parametrize("precision", [100.0, 1000.0, 10000.0])
def test_stored_precision_is_reported(tmp_path, precision):
    # the precision read_xtc returns (calcIR.cu:342-345 rounds the frame
    # spacing with it); frames of <= 9 atoms are stored without one
    xyz, box, times = C.synth_trajectory(8, 3, 2)
    f = tmp_path / "p.xtc"
    C.xtc_write(f, xyz, box, times, precision=precision)
    assert C.xtc_precision(f, 0) == precision and C.xtc_precision(f, 1) == precision
    with pytest.raises(C.CalcirError):
        C.xtc_precision(f, 3)
    g = tmp_path / "small.xtc"
    C.xtc_write(g, xyz[:, :8], box, times, precision=precision)
    assert C.xtc_precision(g, 1) == -1.0


class _Bits:
    """Bit stream as the xdr3dfcoord code writes it: each field's bits most
    significant first, packed into bytes from their top bit."""

    def __init__(self):
        self.bits = []

    def put(self, nbits, value):
        for k in range(nbits - 1, -1, -1):
            self.bits.append((value >> k) & 1)

    def bytes(self):
        b = self.bits + [0] * (-len(self.bits) % 8)
        return bytes(int("".join(map(str, b[i:i + 8])), 2) for i in range(0, len(b), 8))


def _pack_ints(bits, nbits, sizes, nums):
    """The published mixed-radix packing of a triple: the number
    (n0 * s1 + n1) * s2 + n2, its bytes least significant first, 8 bits each
    (most significant bit first), the last byte with the remaining
    nbits - 8 (nbytes - 1) bits."""
    v = (nums[0] * sizes[1] + nums[1]) * sizes[2] + nums[2]
    nbytes = max(1, (v.bit_length() + 7) // 8)
    by = [(v >> (8 * k)) & 0xFF for k in range(nbytes)]
    if nbits >= 8 * nbytes:
        for x in by:
            bits.put(8, x)
        bits.put(nbits - 8 * nbytes, 0)
    else:
        for x in by[:-1]:
            bits.put(8, x)
        bits.put(nbits - 8 * (nbytes - 1), by[-1])


def _bits_of_product(sizes):
    """Bits of a packed triple: the bit length of s0 s1 s2 (the format's
    count, one more than needed when the product is a power of two)."""
    return (sizes[0] * sizes[1] * sizes[2]).bit_length()


def test_hand_assembled_compressed_frame(tmp_path):
    # A 10-atom frame at precision 1000 (the compressed path: more than 9
    # atoms), assembled bit by bit from the published xdr3dfcoord layout
    # rather than by this project's writer.  Atoms 0..5 are coded as plain
    # offsets from minint (one mixed-radix triple each), the first with a
    # run flag (1, then run + is_smaller + 1 = 1 in 5 bits), the others with
    # flag 0; atom 6 is a plain triple with a run of two small atoms (flag 1,
    # value 7), each small triple in smallidx bits relative to the previous
    # atom with offset smallnum = magicints[smallidx] / 2, the first one
    # exchanged with the plain atom (the decoder emits it first); atom 9 is a
    # plain triple with flag 1, value 1 (run 0 again).
    prec = 1000.0
    plain = [(1234, -567, 890), (3000, 2500, -100), (-1500, 40, 2222), (777, 3333, 1500), (0, 0, 0),
             (2999, -999, 1111), (500, 600, 700), (-1200, 3100, -50)]
    smallidx = 9          # magicints[9] = 8: triples in [0, 8)^3 packed in 9 bits
    smallnum = 4
    small = [(1, 7, 3), (6, 0, 5)]  # the two small triples after atom 6
    lint = [np.array(c) for c in plain]
    minint = np.min(lint, axis=0)
    maxint = np.max(lint, axis=0)
    sizes = [int(x) for x in (maxint - minint + 1)]
    bitsize = _bits_of_product(sizes)
    bits = _Bits()

    def put_plain(c, flag_value):
        _pack_ints(bits, bitsize, sizes, [int(x) for x in (np.array(c) - minint)])
        if flag_value is None:
            bits.put(1, 0)
        else:
            bits.put(1, 1)
            bits.put(5, flag_value)

    put_plain(plain[0], 1)
    for c in plain[1:6]:
        put_plain(c, None)
    put_plain(plain[6], 7)
    for t in small:
        _pack_ints(bits, smallidx, [8, 8, 8], list(t))
    put_plain(plain[7], 1)
    payload = bits.bytes()
    # expected decoding: atoms 0..5 plain; then S1 = small0 + atom6 - smallnum
    # emitted before atom 6; S2 = small1 + S1 - smallnum; then the last plain
    s1 = np.array(small[0]) + np.array(plain[6]) - smallnum
    s2 = np.array(small[1]) + s1 - smallnum
    expect = [np.array(c) for c in plain[:6]] + [s1, np.array(plain[6]), s2, np.array(plain[7])]

    frame = struct.pack(">iiif", 1995, 10, 7, 0.5) + struct.pack(">9f", 3.0, 0, 0, 0, 3.5, 0, 0, 0, 4.0)
    frame += struct.pack(">i", 10) + struct.pack(">f", prec)
    frame += struct.pack(">3i", *[int(x) for x in minint]) + struct.pack(">3i", *[int(x) for x in maxint])
    frame += struct.pack(">i", smallidx) + struct.pack(">i", len(payload))
    frame += payload + b"\0" * (-len(payload) % 4)
    f = tmp_path / "hand.xtc"
    f.write_bytes(frame + frame)  # two identical frames
    assert C.xtc_scan(f) == (10, 2)
    xyz, box, times, steps = C.xtc_read(f)
    inv = np.float32(1.0 / prec)
    want = (np.array(expect, np.int64).astype(np.float32) * inv).astype(np.float32)
    assert np.array_equal(xyz[0], want), (xyz[0] / inv, np.array(expect))
    assert np.array_equal(xyz[1], want)
    assert list(steps) == [7, 7] and np.array_equal(box[0], np.float32([3.0, 3.5, 4.0]))
    assert C.xtc_precision(f, 0) == prec


# magicints of the published xdr3dfcoord code from index 9 (FIRSTIDX) on:
# triples of values below magicints[i] pack into i bits
_MAGIC = {9: 8, 10: 10, 11: 12, 12: 16, 13: 20}


def _compressed_frame(natoms, prec, minint, maxint, smallidx, payload):
    frame = struct.pack(">iiif", 1995, natoms, 3, 0.25) + struct.pack(">9f", 5.0, 0, 0, 0, 5.0, 0, 0, 0, 5.0)
    frame += struct.pack(">i", natoms) + struct.pack(">f", prec)
    frame += struct.pack(">3i", *[int(x) for x in minint]) + struct.pack(">3i", *[int(x) for x in maxint])
    frame += struct.pack(">i", smallidx) + struct.pack(">i", len(payload))
    return frame + payload + b"\0" * (-len(payload) % 4)


class _Coder:
    """Plain atoms and runs of small atoms in the xdr3dfcoord bit layout,
    with the decoder's expected output: a flag bit 0 keeps the previous run
    length (3 x small atoms; 0 at the start), a flag bit 1 is followed by 5
    bits of run + is_smaller + 1 (is_smaller in {-1, 0, 1}); the small atoms
    of a run use the precision index in force before the flag's change; the
    first small atom is emitted before its plain atom; after the flag the
    index moves by is_smaller and the offset smallnum becomes
    magicints[index] / 2 (up) or the previous 'smaller' (down)."""

    def __init__(self, minint, maxint, smallidx, large):
        self.bits = _Bits()
        self.minint = np.array(minint, np.int64)
        self.sizes = [int(x) for x in np.array(maxint, np.int64) - self.minint + 1]
        self.large = large
        self.bitsize = _bits_of_product(self.sizes)
        self.axbits = [s.bit_length() for s in self.sizes]
        self.idx = smallidx
        self.smallnum = _MAGIC[smallidx] // 2
        self.smaller = _MAGIC[max(9, smallidx - 1)] // 2
        self.run = 0
        self.out = []

    def atom(self, c, small=(), is_smaller=0):
        c = np.array(c, np.int64)
        off = [int(x) for x in c - self.minint]
        if self.large:
            for k in range(3):
                self.bits.put(self.axbits[k], off[k])
        else:
            _pack_ints(self.bits, self.bitsize, self.sizes, off)
        run = 3 * len(small)
        if run == self.run and is_smaller == 0:
            self.bits.put(1, 0)
        else:
            self.bits.put(1, 1)
            self.bits.put(5, run + is_smaller + 1)
        self.run = run
        prev = c
        emitted = []
        for k, t in enumerate(small):
            assert all(0 <= v < _MAGIC[self.idx] for v in t)
            _pack_ints(self.bits, self.idx, [_MAGIC[self.idx]] * 3, list(t))
            s = np.array(t, np.int64) + prev - self.smallnum
            emitted += [s, c] if k == 0 else [s]
            prev = s
        self.out += emitted if small else [c]
        self.idx += is_smaller
        if is_smaller < 0:
            self.smallnum = self.smaller
            self.smaller = _MAGIC[self.idx - 1] // 2 if self.idx > 9 else 0
        elif is_smaller > 0:
            self.smaller = self.smallnum
            self.smallnum = _MAGIC[self.idx] // 2


def _check_decoded(tmp_path, name, natoms, prec, minint, maxint, smallidx, coder):
    f = tmp_path / name
    f.write_bytes(_compressed_frame(natoms, prec, minint, maxint, smallidx, coder.bits.bytes()))
    assert C.xtc_scan(f) == (natoms, 1)
    xyz = C.xtc_read(f)[0]
    assert len(coder.out) == natoms
    inv = np.float32(1.0 / prec)
    want = (np.array(coder.out, np.int64).astype(np.float32) * inv).astype(np.float32)
    assert np.array_equal(xyz[0], want), (xyz[0] / inv, np.array(coder.out))


def test_hand_assembled_frame_with_changing_small_precision(tmp_path):
    # runs at precision index 10, then 11 (the first flag raises it), then
    # 10 again (the second lowers it), each run decoded at the index in force
    # before its flag; the third atom's run repeats the second's length under
    # a flag bit 0, the fourth's flag ends the runs; 10 atoms at precision
    # 1000
    minint, maxint = (-2000, -1500, -1000), (3000, 3500, 4000)
    cd = _Coder(minint, maxint, 10, large=False)
    cd.atom((100, 200, 300), small=[(1, 9, 4), (7, 0, 8)], is_smaller=1)
    cd.atom((-1900, 3400, 3900), small=[(11, 0, 6)], is_smaller=-1)
    cd.atom((2999, -1499, 17), small=[(0, 9, 9)])
    cd.atom((0, 0, 0))
    cd.atom((-2000, -1500, -1000))
    cd.atom((3000, 3500, 4000))
    _check_decoded(tmp_path, "adapt.xtc", 10, 1000.0, minint, maxint, 10, cd)


def test_hand_assembled_frame_on_the_large_range_path(tmp_path):
    # an axis spanning more than 0xffffff integers codes every plain atom
    # axis by axis (bit length of each axis' size) instead of one mixed-radix
    # triple; runs and their small triples are as on the normal path.  11
    # atoms at precision 1e6, integers within 2^24 so that they are exact
    # in FP32
    minint, maxint = (-9000000, -100, 5), (9000000, 250000, 70000)
    cd = _Coder(minint, maxint, 12, large=True)
    cd.atom((1234567, 99, 60000), small=[(15, 0, 8)])
    for c in [(-9000000, -100, 5), (9000000, 250000, 70000), (0, 0, 6), (-42, 1717, 999),
              (8388607, -99, 12345), (-8388607, 123456, 54321)]:
        cd.atom(c)
    cd.atom((5, 5, 5), is_smaller=-1)   # a flag with no run that lowers the index to 11
    cd.atom((700, 800, 900), small=[(11, 3, 0)], is_smaller=-1)
    _check_decoded(tmp_path, "large.xtc", 11, 1e6, minint, maxint, 12, cd)


_FUZZ = r"""
import os, sys
import numpy as np
import calcir_amd as C
d, natoms, iters, seed = sys.argv[1], int(sys.argv[2]), int(sys.argv[3]), int(sys.argv[4])
rng = np.random.default_rng(seed)
xyz, box, times = C.synth_trajectory((natoms + 3) // 4, 3, seed)
xyz = xyz[:, :natoms]
good_p, bad_p = os.path.join(d, "good.xtc"), os.path.join(d, "bad.xtc")
C.xtc_write(good_p, xyz, box, times)
good = np.frombuffer(open(good_p, "rb").read(), np.uint8)
refused = read = 0
for it in range(iters):
    b = good.copy()
    kind = it % 4
    if kind == 0:
        for _ in range(rng.integers(1, 4)):
            b[rng.integers(0, b.size)] ^= np.uint8(1 << rng.integers(0, 8))
    elif kind == 1:
        b = b[: rng.integers(0, b.size)]
    elif kind == 2:
        p = rng.integers(0, b.size - 4)
        b[p:p + 4] = rng.integers(0, 256, 4)
    else:
        p = rng.integers(60, b.size)
        b[p:] = rng.integers(0, 256, b.size - p)
    open(bad_p, "wb").write(b.tobytes())
    try:
        na, nf = C.xtc_scan(bad_p)
        if na == natoms and 1 <= nf <= 3:
            C.xtc_read(bad_p)
            C.xtc_precision(bad_p, 0)
            read += 1
    except C.CalcirError:
        refused += 1
print("done", iters, refused, read, flush=True)
"""


@pytest.mark.parametrize("natoms", [9, 10, 160])
def test_mutated_files_are_read_or_refused(tmp_path, natoms):
    # bit flips, truncations, overwritten words and random tails of a valid
    # file: every call returns or raises CalcirError, none reads out of
    # bounds or crashes (run in a child process, so a crash fails this test
    # instead of ending the session)
    r = subprocess.run([sys.executable, "-c", _FUZZ, str(tmp_path), str(natoms), "600", str(natoms)],
                       capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, PYTHONPATH=os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    assert r.returncode == 0, r.stderr[-2000:]
    done, iters, refused, read = r.stdout.split()
    assert (done, iters) == ("done", "600") and int(refused) > 50 and int(read) > 50, r.stdout
