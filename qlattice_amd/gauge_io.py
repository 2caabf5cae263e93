"""NERSC gauge configuration I/O (qlat qcd.h:34-270; gf_avg_plaq and
gf_avg_link_trace, qcd.cpp:84-240; unitarize, qlat-utils matrix-hmc.h:29-74).

File layout: a text header between ``BEGIN_HEADER`` and ``END_HEADER``, then,
as the last bytes of the file, the links of every site in reference order
(x fastest), mu = 0..3, as 6 (``4D_SU3_GAUGE``: rows 0 and 1) or 9
(``4D_SU3_GAUGE_3x3``) complex numbers per link, IEEE754 doubles big- or
little-endian.  The header's CHECKSUM is the sum modulo 2^32 of the data's
32-bit words in little-endian (machine) order, CRC32HASH the zlib CRC-32 of
the data block as stored (field_simple_checksum, field_crc32;
qlat field-io.h:11-27, :84-87).  Two-row links get their third row from
``unitarize``.  A checksum mismatch is a warning, as in the reference.
"""

import os
import time
import warnings
import zlib

import numpy as np

from .fields import Geometry


class GaugeFieldInfo:
    """GaugeFieldInfo (qcd.h:34-64) with the reference's defaults."""

    def __init__(self):
        self.ensemble_id = "42"
        self.ensemble_label = "default-ensemble"
        self.creator = "Qlat"
        self.date = time.ctime() + "\n"
        self.datatype = "4D_SU3_GAUGE"
        self.floating_point = "IEEE64BIG"
        self.sequence_num = 0
        self.beta = 0.0
        self.plaq = 1.0
        self.trace = 0.0
        self.simple_checksum = 0
        self.crc32 = 0
        self.total_site = (0, 0, 0, 0)


def make_gauge_field_header(gfi=None):
    """qcd.h:66-91."""
    gfi = gfi or GaugeFieldInfo()
    lines = [
        "BEGIN_HEADER\n",
        "HDR_VERSION = 1.0\n",
        "DATATYPE = 4D_SU3_GAUGE\n",
    ]
    lines += [f"DIMENSION_{m + 1} = {int(gfi.total_site[m])}\n" for m in range(4)]
    lines += [
        f"LINK_TRACE = {gfi.trace:.12f}\n",
        f"PLAQUETTE = {gfi.plaq:.12f}\n",
        f"CHECKSUM = {gfi.simple_checksum:08x}\n",
        f"CRC32HASH = {gfi.crc32:08x}\n",
        f"CREATOR = {gfi.creator}\n",
        f"ARCHIVE_DATE = {gfi.date}",  # ctime's string ends with a newline
        f"ENSEMBLE_ID = {gfi.ensemble_id}\n",
        f"ENSEMBLE_LABEL = {gfi.ensemble_label}\n",
        f"BETA = {gfi.beta:.12f}\n",
        f"SEQUENCE_NUMBER = {int(gfi.sequence_num)}\n",
        "FLOATING_POINT = IEEE64BIG\n",
        "END_HEADER\n",
    ]
    return "".join(lines)


def read_gauge_field_header(path):
    """qcd.h:93-141: the header fields the loader uses (None if the file does
    not start with BEGIN_HEADER)."""
    gfi = GaugeFieldInfo()
    with open(path, "rb") as f:
        if f.readline() != b"BEGIN_HEADER\n":
            return None
        props = {}
        for raw in f:
            line = raw.decode("ascii", errors="replace")
            if line == "END_HEADER\n" or line == "":
                break
            if " = " in line:
                k, v = line.split(" = ", 1)
                props[k] = v.rstrip("\n")
    gfi.total_site = tuple(int(props.get(f"DIMENSION_{m + 1}", "0")) for m in range(4))
    gfi.trace = float(props.get("LINK_TRACE", "0"))
    gfi.plaq = float(props.get("PLAQUETTE", "0"))
    gfi.simple_checksum = int(props["CHECKSUM"], 16) if "CHECKSUM" in props else 0
    gfi.crc32 = int(props["CRC32HASH"], 16) if "CRC32HASH" in props else 0
    gfi.datatype = props.get("DATATYPE", "")
    gfi.floating_point = props.get("FLOATING_POINT", "")
    return gfi


def unitarize(cm):
    """matrix-hmc.h:65-74 on (..., 3, 3): normalise row 0, orthogonalise row 1
    against it and normalise, row 2 = conj(row0 x row1)."""
    cm = np.array(cm, dtype=np.complex128, copy=True)
    p1, p2 = cm[..., 0, :], cm[..., 1, :]
    p1 /= np.sqrt(np.sum(np.abs(p1) ** 2, axis=-1))[..., None]
    c = np.sum(np.conj(p1) * p2, axis=-1)
    p2 -= c[..., None] * p1
    p2 /= np.sqrt(np.sum(np.abs(p2) ** 2, axis=-1))[..., None]
    cm[..., 2, :] = np.conj(np.cross(p1, p2))
    return cm


def gf_avg_plaq(gf):
    """qcd.cpp:84-130: average over sites and the 6 planes of
    Re tr[U_m1(x) U_m2(x+m1) (U_m2(x) U_m1(x+m2))^+] / 3."""
    from .gauge import plaquette

    return plaquette(gf.data, gf.geo.total_site())


def gf_avg_link_trace(gf):
    """qcd.cpp:205-228: average of Re tr U / 3 over all links."""
    return float(np.trace(gf.data, axis1=-2, axis2=-1).real.sum() / 3 / (4 * gf.geo.total_volume()))


def field_simple_checksum(data):
    """Sum of the 32-bit words of the data (little-endian) modulo 2^32."""
    words = np.ascontiguousarray(data).view("<u4")
    return int(np.sum(words, dtype=np.uint64) & 0xFFFFFFFF)


def save_gauge_field(gf, path, gfi=None):
    """qcd.h:143-179: rows 0 and 1 of every link, big-endian doubles, written
    to path + '.partial' and renamed; returns the data size in bytes."""
    gfi = gfi or GaugeFieldInfo()
    rows = np.ascontiguousarray(gf.data[..., :2, :], dtype="<c16").reshape(-1)
    gfi.simple_checksum = field_simple_checksum(rows)
    data = rows.astype(">c16").tobytes()
    gfi.plaq = gf_avg_plaq(gf)
    gfi.trace = gf_avg_link_trace(gf)
    gfi.crc32 = zlib.crc32(data) & 0xFFFFFFFF
    gfi.total_site = gf.geo.total_site()
    with open(path + ".partial", "wb") as f:
        f.write(make_gauge_field_header(gfi).encode("ascii"))
        f.write(data)
    os.replace(path + ".partial", path)
    return len(data)


def load_gauge_field(gf, path):
    """qcd.h:181-270: fills gf (its geometry from the header) and returns the
    data size in bytes (0 if the file has no data)."""
    gfi = read_gauge_field_header(path)
    if gfi is None:
        raise ValueError(f"load_gauge_field: '{path}' has no NERSC header")
    if gfi.datatype == "4D_SU3_GAUGE":
        n = 6
    elif gfi.datatype == "4D_SU3_GAUGE_3x3":
        n = 9
    else:
        raise ValueError(f"load_gauge_field: datatype '{gfi.datatype}'")
    if gfi.floating_point == "IEEE64BIG":
        dt = ">c16"
    elif gfi.floating_point == "IEEE64LITTLE":
        dt = "<c16"
    else:
        raise ValueError(f"load_gauge_field: floating point '{gfi.floating_point}'")
    geo = Geometry(gfi.total_site)
    V = geo.total_volume()
    size = V * 4 * n * 16
    with open(path, "rb") as f:
        f.seek(0, 2)
        if f.tell() < size:
            return 0
        f.seek(-size, 2)
        raw = f.read(size)
    crc = zlib.crc32(raw) & 0xFFFFFFFF
    if crc != gfi.crc32:
        warnings.warn(f"load_gauge_field: '{path}' CHECKSUM= {crc:08X} (calc) {gfi.crc32:08X} (read) "
                      "possibly missing CRC32HASH field")
    vals = np.frombuffer(raw, dtype=dt).astype("<c16")
    simple = field_simple_checksum(vals)
    if simple != gfi.simple_checksum:
        warnings.warn(f"load_gauge_field: '{path}' CHECKSUM= {simple:08X} (calc) {gfi.simple_checksum:08X} (read)")
    links = vals.reshape(V, 4, n)
    U = np.zeros((V, 4, 3, 3), dtype=np.complex128)
    U[:, :, 0, :] = links[:, :, 0:3]
    U[:, :, 1, :] = links[:, :, 3:6]
    if n == 9:
        U[:, :, 2, :] = links[:, :, 6:9]
    else:
        U = unitarize(U)
    gf.geo = geo
    gf.data = U
    return size
