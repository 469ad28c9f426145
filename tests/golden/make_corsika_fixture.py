"""Writes tests/golden/muon_ring4_reference.npz: what the reference's own
CORSIKA IACT reader (its src/bernlohr/ eventio library, built by
oracle/build_ref.sh into oracle/_ref/corsika_dump) reads from the tutorial
file muon_ring4.corsika.gz (tutorials/ of the reference, kept here as a data
fixture): the telescope positions, the event's wavelength bounds and every
photon bunch. Run from the repository root after oracle/build_ref.sh:

    python tests/golden/make_corsika_fixture.py
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def dump(path):
    """Parse corsika_dump's text output into arrays."""
    exe = os.path.join(ROOT, "oracle", "_ref", "corsika_dump")
    out = subprocess.run([exe, path], check=True, capture_output=True, text=True).stdout
    tel, events, bunches = [], [], []
    for line in out.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "P":
            tel.append([float(v) for v in f[2:6]])
        elif f[0] == "E":
            events.append([float(v) for v in f[1:4]])
        elif f[0] == "B":
            bunches.append([float(v) for v in f[1:]])
    b = np.array(bunches).reshape(-1, 11)
    return dict(telescopes=np.array(tel).reshape(-1, 4), events=np.array(events).reshape(-1, 3),
                array_no=b[:, 0].astype(np.int32), tel_no=b[:, 1].astype(np.int32),
                x=b[:, 2].astype(np.float32), y=b[:, 3].astype(np.float32), zem=b[:, 4].astype(np.float32),
                time=b[:, 5].astype(np.float32), cx=b[:, 6].astype(np.float32), cy=b[:, 7].astype(np.float32),
                cz=b[:, 8].astype(np.float32), lambda_=b[:, 9].astype(np.float32),
                photons=b[:, 10].astype(np.float32))


if __name__ == "__main__":
    d = dump(os.path.join(HERE, "muon_ring4.corsika.gz"))
    np.savez_compressed(os.path.join(HERE, "muon_ring4_reference.npz"), **d)
    print(f"{len(d['x'])} bunches, {len(d['telescopes'])} telescopes, events {d['events'].tolist()}", file=sys.stderr)
