"""ACorsikaIACTFile (include/ACorsikaIACTFile.h, src/ACorsikaIACTFile.cxx):
CORSIKA IACT output files read into the photon-bunch table that
rbgpu_bunches_to_rays turns into rays on the device.

The files are eventio streams (optionally gzip-compressed). Every top-level
item starts with the sync tag 0xD41F8A37, whose byte order gives the order of
the item, then a type/version word (type in bits 0-15, version in bits
20-31), an ident word, and a length word (bits 0-29; bit 30: the item holds
only sub-items; bit 31: an extension word follows whose bits 0-11 are length
bits 30-41). Sub-items carry the same header without the sync tag. Scalars
are int16 ("short"), int32 ("long") and float32 ("real").

Items read, as ACorsikaIACTFile::Open and ::ReadEvent do (:190-470): the run
header (1200), input card (1212), telescope positions (1201: n, then x[n],
y[n], z[n], r[n]), event header (1202: the CORSIKA EVTH block; words 95 and
96 bound the wavelength), telescope offsets (1203), the telescope array
(1204, with a photon-bunch sub-item 1205 per telescope, or the bunch items
between 1213 and 1214), and the event end (1209). A photon-bunch item of
version 0 holds 8 floats per bunch (x, y, cx, cy, ctime, zem, photons,
lambda); version 1000 holds 8 int16 per bunch, scaled as the eventio reader
scales them (x, y, ctime x 0.1; cx, cy / 30000 clipped to [-1, 1]; zem =
10^(0.001 s); photons x 0.01; lambda as is). cz = -sqrt(1 - cx^2 - cy^2) with
the squares in float, as the reader computes it (:452).
"""
import gzip
import struct

import numpy as np

from .sources import CorsikaBunches, GetRayArray as _get_ray_array

SYNC = 0xD41F8A37
RUNH, TELPOS, EVTH, TELOFF, TELARRAY, PHOTONS = 1200, 1201, 1202, 1203, 1204, 1205
EVTE, RUNE, LONGI, INPUTCFG, TELARRAY_HEAD, TELARRAY_END = 1209, 1210, 1211, 1212, 1213, 1214
K_MAX_ARRAYS, K_MAX_TELESCOPES = 100, 1000


class _Item:
    __slots__ = ("type", "version", "ident", "data", "order")

    def __init__(self, typ, version, ident, data, order):
        self.type, self.version, self.ident, self.data, self.order = typ, version, ident, data, order


def _header(buf, pos, order):
    """type/version, ident, length (+ extension) at buf[pos:]; returns the
    item fields and the offset of its data."""
    tv, ident, length = struct.unpack_from(order + "IiI", buf, pos)
    pos += 12
    if length & 0x80000000:
        (ext,) = struct.unpack_from(order + "I", buf, pos)
        pos += 4
        length = (length & 0x3FFFFFFF) | ((ext & 0xFFF) << 30)
    else:
        length &= 0x3FFFFFFF
    return tv & 0xFFFF, (tv >> 20) & 0xFFF, ident, length, pos


def _top_items(raw):
    """Top-level items of an eventio byte stream, skipping to the next sync
    tag when the stream is out of step (find_io_block)."""
    pos, n = 0, len(raw)
    le, be = struct.pack("<I", SYNC), struct.pack(">I", SYNC)
    while pos + 16 <= n:
        tag = raw[pos:pos + 4]
        if tag not in (le, be):
            a, b = raw.find(le, pos + 1), raw.find(be, pos + 1)
            cands = [c for c in (a, b) if c >= 0]
            if not cands:
                return
            pos = min(cands)
            continue
        order = "<" if tag == le else ">"
        typ, ver, ident, length, start = _header(raw, pos + 4, order)
        if start + length > n:
            return
        yield _Item(typ, ver, ident, memoryview(raw)[start:start + length], order)
        pos = start + length


def _sub_items(item):
    buf, pos, n = item.data, 0, len(item.data)
    while pos + 12 <= n:
        typ, ver, ident, length, start = _header(buf, pos, item.order)
        if start + length > n:
            return
        yield _Item(typ, ver, ident, buf[start:start + length], item.order)
        pos = start + length


def _tel_block(item):
    """read_tel_block: a count, one int32, then count - 1 floats."""
    o = item.order
    (length,) = struct.unpack_from(o + "i", item.data, 0)
    (first,) = struct.unpack_from(o + "i", item.data, 4)
    vals = np.frombuffer(item.data, dtype=o + "f4", count=max(length - 1, 0), offset=8).astype(np.float64)
    return np.concatenate([[float(first)], vals])


def _photons(item):
    """read_tel_photons: (array, tel, bunch columns) or None for an item
    version the reader refuses."""
    if item.version % 1000 != 0:
        return None
    o = item.order
    arr, tel, _photons_sum, nb = struct.unpack_from(o + "hhfi", item.data, 0)
    if item.version // 1000 == 0:
        b = np.frombuffer(item.data, dtype=o + "f4", count=8 * nb, offset=12).reshape(nb, 8)
        x, y, cx, cy, ctime, zem, photons, lam = (b[:, k].astype(np.float32) for k in range(8))
    elif item.version // 1000 == 1:
        s = np.frombuffer(item.data, dtype=o + "i2", count=8 * nb, offset=12).reshape(nb, 8).astype(np.float64)
        f32 = lambda v: v.astype(np.float32)  # noqa: E731
        x, y = f32(0.1 * s[:, 0]), f32(0.1 * s[:, 1])
        cx, cy = f32(np.clip(s[:, 2] / 30000.0, -1.0, 1.0)), f32(np.clip(s[:, 3] / 30000.0, -1.0, 1.0))
        ctime, zem = f32(0.1 * s[:, 4]), f32(np.power(10.0, 0.001 * s[:, 5]))
        photons, lam = f32(0.01 * s[:, 6]), f32(s[:, 7])
    else:
        return None
    return arr, tel, dict(x=x, y=y, cx=cx, cy=cy, time=ctime, zem=zem, photons=photons, lambda_=lam)


class ACorsikaIACTFile:
    """The reader side of ACorsikaIACTFile: Open, ReadEvent, the telescope
    positions and wavelength bounds, GetBunches and GetRayArray."""

    def __init__(self):
        self.Close()
        self.fMaxPhotonBunches = 100000

    def Close(self):
        self._items = None
        self.fNumberOfTelescopes = 0
        self.fTelescopePosition = np.zeros((4, 0))
        self.fRunHeader = None
        self.fEventHeader = None
        self.fBunches = None
        self.fMinWavelength = self.fMaxWavelength = None

    def IsOpen(self):
        return self._items is not None

    def Open(self, fname):
        """Reads the run header, input card and telescope positions (:190-260);
        raises if the file does not start with them."""
        with open(fname, "rb") as f:
            raw = f.read()
        if raw[:2] == b"\x1f\x8b":
            raw = gzip.decompress(raw)
        self._items = _top_items(raw)
        first = next(self._items, None)
        if first is None or first.type != RUNH:
            self.Close()
            raise ValueError(f"{fname}: the first item is not the run header (1200)")
        self.fRunHeader = _tel_block(first)
        second = next(self._items, None)
        if second is None or second.type != INPUTCFG:
            self.Close()
            raise ValueError(f"{fname}: the second item is not the input card (1212)")
        third = next(self._items, None)
        if third is None or third.type != TELPOS:
            self.Close()
            raise ValueError(f"{fname}: the third item is not the telescope positions (1201)")
        (ntel,) = struct.unpack_from(third.order + "i", third.data, 0)
        ntel = min(ntel, K_MAX_TELESCOPES)
        pos = np.frombuffer(third.data, dtype=third.order + "f4", count=4 * ntel, offset=4)
        self.fNumberOfTelescopes = ntel
        self.fTelescopePosition = pos.reshape(4, ntel).astype(np.float64)

    def ReadEvent(self, num):
        """Reads event `num` (from 1) into the bunch table (:290-470); returns
        num, or -1 when the file ends first or the event was passed."""
        if not self.IsOpen():
            return -1
        if self.fEventHeader is not None:
            if int(self.fEventHeader[1]) == num:
                return num
            if int(self.fEventHeader[1]) > num:
                return -1
        seen = set()
        cols = {k: [] for k in ("tel_no", "array_no", "x", "y", "zem", "time", "cx", "cy", "lambda_", "photons")}

        def add(arr, tel, b):
            cols["tel_no"].append(np.full(len(b["x"]), tel, np.int32))
            cols["array_no"].append(np.full(len(b["x"]), arr, np.int32))
            for k in ("x", "y", "zem", "time", "cx", "cy", "lambda_", "photons"):
                cols[k].append(b[k])

        individual = False
        for item in self._items:
            t = item.type
            if t == EVTH:
                seen = set()
                evth = _tel_block(item)
                if int(evth[1]) != num:
                    continue
                self.fEventHeader = evth
                self.fMaxWavelength, self.fMinWavelength = float(evth[96]), float(evth[95])
                cols = {k: [] for k in cols}
                seen.add(EVTH)
            elif t == TELOFF and EVTH in seen:
                seen.add(TELOFF)
            elif t in (TELARRAY, TELARRAY_HEAD) and EVTH in seen and TELOFF in seen:
                seen.add(t)
                if t == TELARRAY:
                    photons = [s for s in _sub_items(item) if s.type == PHOTONS][:self.fNumberOfTelescopes]
                    for s in photons:
                        got = _photons(s)
                        if got is not None and got[1] >= 0:
                            add(*got)
                else:
                    individual = True
            elif t == PHOTONS and individual:
                got = _photons(item)
                if got is not None and got[1] >= 0:
                    add(*got)
            elif t == TELARRAY_END:
                individual = False
            elif t == EVTE and EVTH in seen:
                seen.add(EVTE)
            if {EVTH, TELOFF, EVTE} <= seen and (TELARRAY in seen or TELARRAY_HEAD in seen):
                cat = {k: (np.concatenate(v) if v else np.zeros(0)) for k, v in cols.items()}
                self.fBunches = CorsikaBunches(cat["tel_no"], cat["array_no"], cat["x"], cat["y"], cat["zem"],
                                               cat["time"], cat["cx"], cat["cy"], cat["lambda_"], cat["photons"])
                return num
        return -1

    def GetBunches(self):
        return self.fBunches

    def GetNumberOfTelescopes(self):
        return self.fNumberOfTelescopes

    def _tel(self, k, i):
        return float(self.fTelescopePosition[k, i]) if 0 <= i < self.fNumberOfTelescopes else float("nan")

    def GetTelescopeX(self, i):
        return self._tel(0, i)

    def GetTelescopeY(self, i):
        return self._tel(1, i)

    def GetTelescopeZ(self, i):
        return self._tel(2, i)

    def GetTelescopeR(self, i):
        return self._tel(3, i)

    def SetMaxPhotonBunches(self, n):
        self.fMaxPhotonBunches = int(n)

    def GetRayArray(self, ctx, telNo, arrayNo, z, refractiveIndex, seed=0):
        """:71-131 on the device: DeviceRays for telescope telNo of array
        arrayNo, starting at height z above the observation level (None for
        indices the reference rejects or before an event is read)."""
        if self.fBunches is None or telNo < 0 or telNo >= self.fNumberOfTelescopes:
            return None
        if arrayNo < 0 or arrayNo >= K_MAX_ARRAYS:
            return None
        return _get_ray_array(ctx, self.fBunches, telNo, arrayNo, z, refractiveIndex, self.GetTelescopeZ(telNo),
                              self.fMinWavelength, self.fMaxPhotonBunches, seed)
