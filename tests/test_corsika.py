"""CORSIKA IACT input: the eventio reader (robast_amd.corsika, mirroring
ACorsikaIACTFile::Open / ReadEvent, src/ACorsikaIACTFile.cxx:190-470) and its
bunch table turned into rays on the device (GetRayArray, :71-131).

CPU (pins the reader): the reference's own tutorial file muon_ring4.corsika.gz
read by the reference's own reader library (src/bernlohr/, built by
oracle/build_ref.sh; its output kept as tests/golden/muon_ring4_reference.npz
by tests/golden/make_corsika_fixture.py) equals this reader's bunch table
field for field; synthetic files in the compact bunch format, in both byte
orders, read the same through both readers when the reference tree is
present; the headers, event selection and error cases; and the conversion of
the file's bunches by the oracle meets the bunch geometry.

GPU: GetRayArray on the file's bunches equals the oracle's conversion bit for
bit."""
import math
import os
import struct

import numpy as np
import pytest

from robast_amd.corsika import ACorsikaIACTFile
from robast_amd import sources as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
FILE = os.path.join(GOLDEN, "muon_ring4.corsika.gz")
REF_READER = os.path.join(ROOT, "oracle", "_ref", "corsika_dump")
FIELDS = ("tel_no", "array_no", "x", "y", "zem", "time", "cx", "cy", "cz", "lambda_", "photons")


def opened(path=FILE, event=1):
    f = ACorsikaIACTFile()
    f.Open(path)
    assert f.ReadEvent(event) == event
    return f


def test_reader_matches_reference_reader():
    ref = np.load(os.path.join(GOLDEN, "muon_ring4_reference.npz"))
    f = opened()
    b = f.GetBunches()
    assert b.n == len(ref["x"]) == 2842
    for k in FIELDS:
        assert np.array_equal(getattr(b, k), ref[k]), k
    assert f.GetNumberOfTelescopes() == len(ref["telescopes"]) == 1
    for i, (x, y, z, r) in enumerate(ref["telescopes"]):
        assert (f.GetTelescopeX(i), f.GetTelescopeY(i), f.GetTelescopeZ(i), f.GetTelescopeR(i)) == (x, y, z, r)
    ev, lmin, lmax = ref["events"][0]
    assert (f.fEventHeader[1], f.fMinWavelength, f.fMaxWavelength) == (ev, lmin, lmax)


def test_reader_headers_and_errors(tmp_path):
    f = opened()
    assert f.ReadEvent(1) == 1                    # already read
    assert f.ReadEvent(0) == -1                   # earlier events cannot be reached
    assert math.isnan(f.GetTelescopeZ(1)) and math.isnan(f.GetTelescopeX(-1))
    assert f.GetRayArray(None, 5, 0, 3000.0, 1.0) is None and f.GetRayArray(None, 0, 100, 3000.0, 1.0) is None
    g = ACorsikaIACTFile()
    g.Open(FILE)
    assert g.ReadEvent(2) == -1                   # the file holds one event
    bad = tmp_path / "bad.dat"
    bad.write_bytes(b"not an eventio file at all")
    with pytest.raises(ValueError):
        ACorsikaIACTFile().Open(str(bad))


# ---- synthetic files in the compact bunch format

def _item(typ, ver, ident, payload, order, top, only_sub=False):
    payload += b"\0" * (-len(payload) % 4)
    length = len(payload) | (0x40000000 if only_sub else 0)
    head = struct.pack(order + "IiI", typ | (ver << 20), ident, length)
    return (struct.pack(order + "I", 0xD41F8A37) if top else b"") + head + payload


def _tel_block(first, vals, order):
    return struct.pack(order + "ii", len(vals) + 1, first) + struct.pack(order + f"{len(vals)}f", *vals)


def synthetic(order, seed):
    rng = np.random.default_rng(seed)
    runh = _tel_block(1, [7.0] * 272, order)
    evth = [0.0] * 272
    evth[0], evth[94], evth[95] = 1.0, 250.0, 700.0   # words 1, 95, 96 of the block
    tels = [(0.0, 0.0, 1200.0, 600.0), (5000.0, -300.0, 1250.0, 700.0)]
    telpos = struct.pack(order + "i", 2) + b"".join(struct.pack(order + "2f", *c) for c in zip(*tels))
    subs, expect = b"", []
    for tel in (0, 1):
        nb = 40
        s = np.stack([rng.integers(-30000, 30000, nb), rng.integers(-30000, 30000, nb),
                      rng.integers(-3000, 3000, nb), rng.integers(-3000, 3000, nb), rng.integers(-500, 500, nb),
                      rng.integers(3000, 6000, nb), rng.integers(1, 800, nb), rng.integers(0, 2, nb) * 420],
                     1).astype(np.int16)
        body = struct.pack(order + "hhfi", 0, tel, float(s[:, 6].sum()) * 0.01, nb) + s.astype(order + "i2").tobytes()
        subs += _item(1205, 1000, tel, body, order, top=False)
        expect.append(s)
    stream = (_item(1200, 0, 1, runh, order, True) + _item(1212, 0, 1, struct.pack(order + "i", 0), order, True)
              + _item(1201, 0, 1, telpos, order, True)
              + _item(1202, 0, 1, _tel_block(1, evth, order), order, True)
              + _item(1203, 0, 1, struct.pack(order + "i", 0), order, True)
              + _item(1204, 0, 0, subs, order, True, only_sub=True)
              + _item(1209, 0, 1, struct.pack(order + "i", 0), order, True))
    return stream, expect


@pytest.mark.parametrize("order", ["<", ">"])
def test_compact_bunches(tmp_path, order):
    stream, shorts = synthetic(order, seed=3 if order == "<" else 4)
    path = tmp_path / f"synthetic{'le' if order == '<' else 'be'}.dat"
    path.write_bytes(stream)
    f = opened(str(path))
    b = f.GetBunches()
    s = np.concatenate(shorts).astype(np.float64)
    assert b.n == 80 and list(b.tel_no) == [0] * 40 + [1] * 40
    assert np.array_equal(b.x, (0.1 * s[:, 0]).astype(np.float32))
    assert np.array_equal(b.cx, np.clip(s[:, 2] / 30000.0, -1, 1).astype(np.float32))
    assert np.array_equal(b.photons, (0.01 * s[:, 6]).astype(np.float32))
    assert np.array_equal(b.lambda_, s[:, 7].astype(np.float32))
    assert (f.fMinWavelength, f.fMaxWavelength, f.GetTelescopeZ(1)) == (250.0, 700.0, 1250.0)
    if not os.path.exists(REF_READER):
        pytest.skip("the reference reader is not built (oracle/build_ref.sh needs the reference tree)")
    import sys
    sys.path.insert(0, GOLDEN)
    from make_corsika_fixture import dump
    ref = dump(str(path))
    for k in FIELDS:
        assert np.array_equal(getattr(b, k), ref[k]), k


def test_oracle_converts_file_bunches():
    from oracle import oracle as O
    f = opened()
    b = f.GetBunches()
    o = S.bunch_opts(0, 0, 30 * 100.0, f.GetTelescopeZ(0), 1.0, f.fMinWavelength, f.fMaxPhotonBunches, seed=5)
    s = O.bunches_to_rays(b, o)
    src = np.repeat(np.arange(b.n), np.ceil(b.photons).astype(int))
    assert len(s["x"]) == len(src) > 10000
    d = np.stack([s["dx"], s["dy"], s["dz"]], 1)
    step = (o.tel_z - o.z) / d[:, 2]
    assert np.allclose(s["x"] + step * d[:, 0], b.x[src], rtol=0, atol=1e-9)
    assert np.allclose(s["y"] + step * d[:, 1], b.y[src], rtol=0, atol=1e-9)
    lam = s["lambda_"] / (1e-9 * 100.0)
    assert lam.min() >= 300.0 * (1 - 1e-12)      # every bunch of this file leaves the wavelength to be drawn


@pytest.mark.gpu
def test_gpu_file_bunches_to_rays(ctx):
    from oracle import oracle as O
    print("CORSIKA muon_ring4: GetRayArray on the device", flush=True)
    f = opened()
    rays = f.GetRayArray(ctx, 0, 0, 30 * 100.0, 1.0, seed=5)
    o = S.bunch_opts(0, 0, 30 * 100.0, f.GetTelescopeZ(0), 1.0, f.fMinWavelength, f.fMaxPhotonBunches, seed=5)
    ref = O.bunches_to_rays(f.GetBunches(), o)
    assert rays.n == len(ref["x"])
    g = {k: rays.t[k][:rays.n].cpu().numpy() for k in rays.t}
    for k in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_", "status", "npoints", "last_node"):
        assert np.array_equal(g[k], ref[k]), k
