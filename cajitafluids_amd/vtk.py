"""The output files of cf_write_output / cf_write_vti (VTK XML ImageData
with raw appended data), written through the library and read back with
numpy: the reader checks a file against what the writer was given, and lets
a user load a step without VisIt or ParaView."""
import ctypes as C
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import _lib


def write_vti(path, q, vel, origin, dx, cycle=0, time=0.0, dt=0.0):
    """One piece from host arrays: q (n2, n1, n0) or (n1, n0); vel a list of
    dim arrays of the same shape.  Uses cf_write_vti (no device)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    dim = len(vel)
    if q.ndim != dim:
        raise ValueError("q must have one axis per velocity component")
    n = (C.c_int * 3)(*(list(q.shape[::-1]) + [1] * (3 - dim)))
    org = (C.c_double * 3)(*(list(origin) + [0.0] * (3 - len(origin))))
    v = np.ascontiguousarray(np.concatenate([np.ravel(a) for a in vel]), dtype=np.float64)
    L = _lib.load()
    rc = L.cf_write_vti(os.fsencode(path), dim, n, org, float(dx), int(cycle), float(time),
                        float(dt), q.ctypes.data_as(_lib._dp), v.ctypes.data_as(_lib._dp))
    if rc != _lib.CF_OK:
        raise _lib.CfError(rc, L.cf_last_error(None).decode())


def read_vti(path):
    """dict: extent, origin, spacing, cycle, time, dtime, quantity
    (n2, n1, n0), velocity (n2, n1, n0, 3)."""
    with open(path, "rb") as f:
        data = f.read()
    mark = data.index(b'<AppendedData encoding="raw">')
    start = data.index(b"_", mark) + 1
    head = data[:mark].decode() + "</VTKFile>"
    root = ET.fromstring(head)
    img = root.find("ImageData")
    ext = [int(x) for x in img.get("WholeExtent").split()]
    out = dict(extent=ext,
               origin=[float(x) for x in img.get("Origin").split()],
               spacing=[float(x) for x in img.get("Spacing").split()])
    for arr in img.find("FieldData"):
        val = arr.text.strip()
        out[arr.get("Name").lower()] = int(val) if arr.get("type") == "Int32" else float(val)
    shape = [max(ext[2 * d + 1] - ext[2 * d], 1) for d in range(3)][::-1]
    for arr in img.find("Piece").find("CellData"):
        off = start + int(arr.get("offset"))
        nbytes = int(np.frombuffer(data, dtype="<u8", count=1, offset=off)[0])
        vals = np.frombuffer(data, dtype="<f8", count=nbytes // 8, offset=off + 8)
        ncomp = int(arr.get("NumberOfComponents"))
        out[arr.get("Name")] = vals.reshape(shape + ([ncomp] if ncomp > 1 else []))
    return out


def read_pvti(path):
    """The master file: its whole extent and the pieces' sources, resolved
    against the master's directory."""
    text = open(path).read()
    root = ET.fromstring(text)
    p = root.find("PImageData")
    base = os.path.dirname(path)
    return dict(extent=[int(x) for x in p.get("WholeExtent").split()],
                pieces=[os.path.join(base, pc.get("Source")) for pc in p.findall("Piece")],
                arrays=[a.get("Name") for a in p.find("PCellData")])


def step_files(directory, step):
    """(master, piece) paths of one step, as cf_write_output names them."""
    return (os.path.join(directory, "CajitaFluids%05d.pvti" % step),
            os.path.join(directory, "raw", "CajitaFluidsOutput%05d%05d.vti" % (0, step)))


__all__ = ["write_vti", "read_vti", "read_pvti", "step_files"]
