"""Extract the golden E/F/S frames from the reference's ASE trajectory.

The reference ships ``gappy/example/ASE-GAPPY/ase.traj`` (ASE 3.18.1, ULM v3
container): 11 NVT-MD frames of a 64-atom carbon diamond cell evaluated by
libgap with the shipped ``gap_parameters`` model.  Each frame stores the
positions, the cell and the raw calculator results (energy in eV, forces in
eV/A, stress in GPa in libgap's order xx yy zz xy yz zx, VASP sign).

ASE is not installed, so this is a small numpy-only reader of the ULM layout:

* bytes 0..23: the tag ``- of UlmASE-Trajectory``;
* bytes 24..47: int64 version (3), number of items, offset of the item table;
* the item table: ``nitems`` int64 offsets;
* at each offset: int64 JSON length, then the JSON text;
* ndarrays inside the JSON are ``{"ndarray": [shape, dtype, absolute_offset]}``.

The output ``c64_diamond/frames.npz`` is data only (numbers, positions, cells,
energies, forces, stresses); it is what the parity tests compare against.

Usage: python tests/golden/extract_traj.py [path/to/ase.traj]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TRAJ = "/root/reference/gappy/example/ASE-GAPPY/ase.traj"


def _ndarray(buf, ref):
    shape, dtype, off = ref["ndarray"]
    count = int(np.prod(shape))
    return np.frombuffer(buf, dtype=np.dtype(dtype).newbyteorder("<"),
                         count=count, offset=off).reshape(shape).copy()


def read_ulm_frames(path):
    buf = open(path, "rb").read()
    if not buf.startswith(b"- of Ulm"):
        raise ValueError("not a ULM file: %s" % path)
    version, nitems, table = np.frombuffer(buf[24:48], dtype="<i8")
    if version != 3:
        raise ValueError("unsupported ULM version %d" % version)
    offsets = np.frombuffer(buf, dtype="<i8", count=int(nitems), offset=int(table))
    frames = []
    numbers = None
    for off in offsets:
        n = int(np.frombuffer(buf, dtype="<i8", count=1, offset=int(off))[0])
        item = json.loads(buf[int(off) + 8:int(off) + 8 + n])
        if "numbers." in item:
            numbers = _ndarray(buf, item["numbers."])
        calc = item["calculator."]
        frames.append(dict(
            numbers=numbers,
            positions=_ndarray(buf, item["positions."]),
            cell=np.array(item["cell"], dtype=np.float64),
            energy=float(calc["energy"]),
            forces=_ndarray(buf, calc["forces."]),
            stress=np.array(calc["stress"], dtype=np.float64),
        ))
    return frames


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_TRAJ
    frames = read_ulm_frames(path)
    out = os.path.join(HERE, "c64_diamond", "frames.npz")
    np.savez_compressed(
        out,
        numbers=frames[0]["numbers"].astype(np.int32),
        positions=np.stack([f["positions"] for f in frames]),
        cells=np.stack([f["cell"] for f in frames]),
        energies=np.array([f["energy"] for f in frames]),
        forces=np.stack([f["forces"] for f in frames]),
        stresses=np.stack([f["stress"] for f in frames]),
    )
    print("wrote %s (%d frames)" % (out, len(frames)))


if __name__ == "__main__":
    main()
