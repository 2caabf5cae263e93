"""Compressed eigen-system files (qlat compressed-eigen-io.h:37-72, :82-195;
commpressed-eigen-io.cpp:16-170, :258-284, :536-646, :748-862, :1005-1055)
and load_low_modes (dslash.cpp:9-23).

A directory ``path`` holds

  metadata.txt      ``s[i] = `` node extents (i < 4) and ``s[4] = `` Ls,
                    ``b[i] = `` block extents, ``nb[i] = `` blocks per node,
                    ``neig``, ``nkeep``, ``nkeep_single``, ``blocks``,
                    ``FP16_COEF_EXP_SHARE_FLOATS``, ``crc32[k] = `` (hex, one
                    per file) and optionally ``gs[i] = `` the lattice extents
  eigen-values.txt  the count, then one value a line
  NN/KKKKKKKKKK.compressed
                    one file per node k (node coordinates x fastest), in
                    directory NN = k / ceil(n_files / min(n_files, 32))

Each file covers the node's blocks (block index x fastest inside the node)
in three regions: the first ``nkeep_single`` basis vectors of every block as
little-endian FP32 complex, the remaining ``nkeep - nkeep_single`` basis
vectors as 16-bit values in groups of 24 floats (one 5D half site) sharing
an exponent word, and for every eigenvector, for every block, its
``nkeep_single`` FP32 coefficients then its 16-bit coefficients in groups of
``FP16_COEF_EXP_SHARE_FLOATS``.  A basis vector is stored with s slowest
(s, half site, spin*3+colour); in memory the half site comes first.  The
header checksum of a file is the CRC-32 of its bytes: the reference XORs
crc32_combine(crc32(chunk), 0, bytes after the chunk) over its chunks, which
is the CRC-32 of their concatenation.

The 16-bit code of a float v in a group of maximum M = sqrt(2)^(e - 32767)
is floor(65536 (v + M) / 2M) (65535 at the top) and decodes to
-M + (code + 1/2) 2M / 65536, computed in FP32 as the reference does
(fp_map, fp_unmap, unmap_fp16_exp).
"""

import math
import os
import zlib

import numpy as np

from .inverter import LowModes

DIST_DIRS = 32  # field-dist-io.h:12
_U16_MAX = 65535
_SQRT2 = 1.4142135623730950488


class CompressedEigenSystemInfo:
    """CompressedEigenSystemInfo (compressed-eigen-io.h:82-117)."""

    def __init__(self, total_site, node_site, block_site, ls, neig, nkeep, nkeep_single,
                 fp16_coef_exp_share_floats, crcs=None):
        self.total_site = tuple(int(v) for v in total_site)
        self.node_site = tuple(int(v) for v in node_site)
        self.block_site = tuple(int(v) for v in block_site)
        self.ls = int(ls)
        self.neig = int(neig)
        self.nkeep = int(nkeep)
        self.nkeep_single = int(nkeep_single)
        self.FP16_COEF_EXP_SHARE_FLOATS = int(fp16_coef_exp_share_floats)
        if not 0 <= self.nkeep_single <= self.nkeep:
            raise ValueError("CompressedEigenSystemInfo: need 0 <= nkeep_single <= nkeep")
        self.nkeep_fp16 = self.nkeep - self.nkeep_single
        for m in range(4):
            if self.total_site[m] % self.node_site[m] or self.node_site[m] % self.block_site[m]:
                raise ValueError(f"CompressedEigenSystemInfo: total_site {self.total_site}, node_site "
                                 f"{self.node_site} and block_site {self.block_site} do not nest")
        if int(np.prod(self.block_site)) % 2:
            raise ValueError("CompressedEigenSystemInfo: the block volume must be even")
        self.total_node = tuple(self.total_site[m] // self.node_site[m] for m in range(4))
        self.total_block = tuple(self.total_site[m] // self.block_site[m] for m in range(4))
        self.node_block = tuple(self.node_site[m] // self.block_site[m] for m in range(4))
        n_files = int(np.prod(self.total_node))
        self.crcs = list(crcs) if crcs is not None else [0] * n_files
        if len(self.crcs) != n_files:
            raise ValueError(f"CompressedEigenSystemInfo: {len(self.crcs)} checksums for {n_files} files")
        nsc = self.FP16_COEF_EXP_SHARE_FLOATS
        if self.nkeep_fp16 and (nsc <= 0 or (2 * self.nkeep_fp16) % nsc):
            raise ValueError("CompressedEigenSystemInfo: FP16_COEF_EXP_SHARE_FLOATS must divide "
                             "2 * (nkeep - nkeep_single)")

    # CompressedEigenSystemData sizes (commpressed-eigen-io.cpp:258-284), in bytes per block
    @property
    def block_vol_eo(self):
        return int(np.prod(self.block_site)) // 2

    @property
    def basis_c_size(self):
        return self.block_vol_eo * self.ls * 12

    def sizes(self):
        nsc = self.FP16_COEF_EXP_SHARE_FLOATS
        basis_single = self.basis_c_size * 8
        basis_fp16 = fp_16_size(self.basis_c_size * 2, 24)
        coef_single = self.nkeep_single * 8
        coef_fp16 = fp_16_size(self.nkeep_fp16 * 2, nsc) if self.nkeep_fp16 else 0
        d = dict(basis_single=basis_single, basis_fp16=basis_fp16, coef_single=coef_single,
                 coef_fp16=coef_fp16, coef_size=coef_single + coef_fp16,
                 bases_size_single=self.nkeep_single * basis_single,
                 bases_size_fp16=self.nkeep_fp16 * basis_fp16)
        d["bases_offset_single"] = 0
        d["bases_offset_fp16"] = d["bases_size_single"]
        d["coefs_offset"] = d["bases_offset_fp16"] + d["bases_size_fp16"]
        d["end_offset"] = d["coefs_offset"] + self.neig * d["coef_size"]
        return d

    def file_size(self):
        return int(np.prod(self.node_block)) * self.sizes()["end_offset"]


def fp_16_size(f_size, nsc):
    """compressed-eigen-io.h:69-72: bytes of f_size floats coded in groups of nsc."""
    return (f_size + f_size // nsc) * 2


def unmap_fp16_exp(e):
    """compressed-eigen-io.h:62-67: sqrt(2)^(e - 32767) rounded to FP32."""
    de = np.asarray(e, dtype=np.int64) - _U16_MAX // 2
    return np.power(_SQRT2, de.astype(np.float32).astype(np.float64)).astype(np.float32)


def fp_unmap(val, vmin, vmax, n=_U16_MAX):
    """compressed-eigen-io.h:56-60 in FP32."""
    f = np.float32
    v = (np.asarray(val, dtype=np.float64) + 0.5).astype(f)
    span = (np.asarray(vmax, dtype=f) - np.asarray(vmin, dtype=f)).astype(f)
    return (np.asarray(vmin, dtype=f) + (v * span).astype(f) / f(n + 1)).astype(f)


def fp_map(v, vmin, vmax, n=_U16_MAX):
    """compressed-eigen-io.h:37-54 in FP32 (the encoder's inverse of fp_unmap)."""
    f = np.float32
    v, vmin, vmax = (np.asarray(a, dtype=f) for a in (v, vmin, vmax))
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (f(n + 1) * ((v - vmin).astype(f) / (vmax - vmin).astype(f)).astype(f)).astype(f)
    t = np.where(vmax > vmin, t, f(0))
    ret = np.trunc(t).astype(np.int64)
    return np.minimum(ret, n)


def read_floats_fp16(words, n, nsc):
    """commpressed-eigen-io.cpp:23-51: n FP32 values from little-endian 16-bit
    words, groups of nsc values after one exponent word each."""
    if n % nsc:
        raise ValueError("read_floats_fp16: n is not a multiple of the group size")
    w = np.asarray(words, dtype=np.uint16).reshape(n // nsc, nsc + 1)
    vmax = unmap_fp16_exp(w[:, 0])[:, None]
    return fp_unmap(w[:, 1:], -vmax, vmax).reshape(n)


def fp16_exponent(vmax):
    """The smallest exponent word e with unmap_fp16_exp(e) >= vmax, so that
    every value of the group lies in [-M, M] (0 for an all-zero group)."""
    vmax = np.asarray(vmax, dtype=np.float32)
    e = np.zeros(vmax.shape, np.int64)
    nz = vmax > 0
    e[nz] = np.ceil(np.log(vmax[nz].astype(np.float64)) / math.log(_SQRT2)).astype(np.int64) + _U16_MAX // 2
    e = np.clip(e, 0, _U16_MAX)
    for _ in range(2):  # correct the float rounding of the logarithm
        low = nz & (unmap_fp16_exp(e) < vmax) & (e < _U16_MAX)
        e[low] += 1
        high = nz & (e > 0) & (unmap_fp16_exp(np.maximum(e - 1, 0)) >= vmax)
        e[high] -= 1
    return e


def write_floats_fp16(vals, nsc):
    """Inverse of read_floats_fp16: 16-bit words of FP32 values, groups of
    nsc values after one exponent word each."""
    v = np.asarray(vals, dtype=np.float32).reshape(-1)
    if v.size % nsc:
        raise ValueError("write_floats_fp16: size is not a multiple of the group size")
    g = v.reshape(-1, nsc)
    e = fp16_exponent(np.abs(g).max(axis=1))
    vmax = unmap_fp16_exp(e)[:, None]
    out = np.empty((g.shape[0], nsc + 1), np.uint16)
    out[:, 0] = e
    out[:, 1:] = fp_map(g, -vmax, vmax)
    return out.reshape(-1)


def dist_file_dir_id(idx, n_files):
    """compute_dist_file_dir_id (field-dist-io.h:63-70)."""
    ndir = min(n_files, DIST_DIRS)
    return idx // ((n_files - 1) // ndir + 1)


def file_name(path, idx, n_files):
    return os.path.join(path, f"{dist_file_dir_id(idx, n_files):02d}", f"{idx:010d}.compressed")


def _get_prop(lines, prop):
    """info_get_prop (qlat-utils show.h:439-448): the rest of the first line
    starting with prop, or ''."""
    for line in lines:
        if line.startswith(prop):
            return line[len(prop):]
    return ""


_GUESSED_SITES = [(64, 64, 64, 128), (48, 48, 48, 96), (48, 48, 48, 64), (24, 24, 24, 64),
                  (32, 32, 32, 64), (16, 16, 16, 32)]


def read_compressed_eigen_system_info(root):
    """commpressed-eigen-io.cpp:82-159 (including the lattice-size guess for
    files without ``gs[i]`` lines)."""
    with open(os.path.join(root, "metadata.txt")) as f:
        lines = f.read().splitlines()

    def geti(prop):
        s = _get_prop(lines, prop)
        if s == "":
            raise ValueError(f"read_compressed_eigen_system_info: '{prop.strip()}' missing in {root}")
        return int(s)

    node_site = [geti(f"s[{i}] = ") for i in range(4)]
    block_site = [geti(f"b[{i}] = ") for i in range(4)]
    for i in range(4):
        if node_site[i] != geti(f"nb[{i}] = ") * block_site[i]:
            raise ValueError(f"read_compressed_eigen_system_info: s[{i}] != nb[{i}] * b[{i}]")
    ls = geti("s[4] = ")
    if geti("b[4] = ") != ls or geti("nb[4] = ") != 1:
        raise ValueError("read_compressed_eigen_system_info: the fifth dimension must be one block")
    neig, nkeep, nkeep_single = geti("neig = "), geti("nkeep = "), geti("nkeep_single = ")
    node_block_vol = int(np.prod([node_site[m] // block_site[m] for m in range(4)]))
    if geti("blocks = ") != node_block_vol:
        raise ValueError("read_compressed_eigen_system_info: 'blocks' does not match s / b")
    nsc = geti("FP16_COEF_EXP_SHARE_FLOATS = ")
    crcs = []
    while True:
        s = _get_prop(lines, f"crc32[{len(crcs)}] = ")
        if s == "":
            break
        crcs.append(int(s.split()[0], 16))
    glb_vol = len(crcs) * int(np.prod(node_site))
    gs = [_get_prop(lines, f"gs[{i}] = ") for i in range(4)]
    if all(s != "" for s in gs):
        total_site = [int(s) for s in gs]
    else:
        total_site = next((list(t) for t in _GUESSED_SITES if int(np.prod(t)) == glb_vol), None)
        if total_site is None:
            total_site = list(node_site)
            total_site[3] *= len(crcs)
    if int(np.prod(total_site)) != glb_vol:
        raise ValueError("read_compressed_eigen_system_info: lattice volume and file count disagree")
    gs5 = _get_prop(lines, "gs[4] = ")
    if gs5 != "" and int(gs5) != ls:
        raise ValueError("read_compressed_eigen_system_info: gs[4] != s[4]")
    return CompressedEigenSystemInfo(total_site, node_site, block_site, ls, neig, nkeep, nkeep_single, nsc,
                                     crcs)


def write_compressed_eigen_system_info(cesi, root):
    """commpressed-eigen-io.cpp:161-197 (metadata.txt.partial, then renamed)."""
    lines = [f"s[{i}] = {cesi.node_site[i]}" for i in range(4)] + [f"s[4] = {cesi.ls}"]
    lines += [f"b[{i}] = {cesi.block_site[i]}" for i in range(4)] + [f"b[4] = {cesi.ls}"]
    lines += [f"nb[{i}] = {cesi.node_block[i]}" for i in range(4)] + ["nb[4] = 1"]
    lines += [f"neig = {cesi.neig}", f"nkeep = {cesi.nkeep}", f"nkeep_single = {cesi.nkeep_single}",
              f"blocks = {int(np.prod(cesi.node_block))}",
              f"FP16_COEF_EXP_SHARE_FLOATS = {cesi.FP16_COEF_EXP_SHARE_FLOATS}"]
    lines += [f"crc32[{i}] = {c:08X}" for i, c in enumerate(cesi.crcs)]
    lines += [f"gs[{i}] = {cesi.total_site[i]}" for i in range(4)] + [f"gs[4] = {cesi.ls}"]
    fn = os.path.join(root, "metadata.txt")
    with open(fn + ".partial", "w") as f:
        f.write("\n".join(lines) + "\n")
    os.replace(fn + ".partial", fn)


def read_eigen_values(path):
    """commpressed-eigen-io.cpp:837-861."""
    with open(os.path.join(path, "eigen-values.txt")) as f:
        lines = f.read().split("\n")
    n = int(lines[0])
    return np.array([float(lines[1 + k]) for k in range(n)], dtype=np.float64)


def write_eigen_values(path, vals):
    """The eigen-values.txt of save_low_modes_decompress (dslash.cpp:125-134)."""
    with open(os.path.join(path, "eigen-values.txt"), "w") as f:
        f.write(f"{len(vals)}\n")
        for v in vals:
            f.write(f"{float(v):.20E}\n")


def _node_block_indices(cesi, idx):
    """Global block indices (x fastest over total_block) of the blocks of
    file idx, in the file's block order (x fastest over node_block)."""
    nc, r = [], idx
    for m in range(4):
        nc.append(r % cesi.total_node[m])
        r //= cesi.total_node[m]
    nb = cesi.node_block
    lc = np.indices(nb[::-1]).reshape(4, -1)[::-1]  # row k: coordinate k, x fastest
    g = [nc[m] * nb[m] + lc[m] for m in range(4)]
    tb = cesi.total_block
    return g[0] + tb[0] * (g[1] + tb[1] * (g[2] + tb[2] * g[3]))


def _decode_file(cesi, raw):
    """Bases (node blocks, nkeep, basis_c_size) and coefs (node blocks, neig,
    nkeep) of one file, in memory order (load_block, commpressed-eigen-io.cpp:748-800)."""
    sz = cesi.sizes()
    nblk = int(np.prod(cesi.node_block))
    bve, ls, ns, nf = cesi.block_vol_eo, cesi.ls, cesi.nkeep_single, cesi.nkeep_fp16
    bases = np.empty((nblk, cesi.nkeep, ls, bve, 12), np.complex64)
    a = nblk * sz["bases_size_single"]
    bases[:, :ns] = np.frombuffer(raw[:a], dtype="<c8").reshape(nblk, ns, ls, bve, 12)
    b = a + nblk * sz["bases_size_fp16"]
    if nf:
        w = np.frombuffer(raw[a:b], dtype="<u2")
        fl = read_floats_fp16(w, nblk * nf * cesi.basis_c_size * 2, 24)
        bases[:, ns:] = fl.view(np.complex64).reshape(nblk, nf, ls, bve, 12)
    bases = np.ascontiguousarray(bases.transpose(0, 1, 3, 2, 4)).reshape(nblk, cesi.nkeep, -1)
    c = np.frombuffer(raw[b:b + cesi.neig * nblk * sz["coef_size"]], dtype=np.uint8)
    c = c.reshape(cesi.neig, nblk, sz["coef_size"])
    coefs = np.empty((nblk, cesi.neig, cesi.nkeep), np.complex64)
    single = np.ascontiguousarray(c[:, :, :sz["coef_single"]]).view("<c8").reshape(cesi.neig, nblk, ns)
    coefs[:, :, :ns] = single.transpose(1, 0, 2)
    if nf:
        w = np.ascontiguousarray(c[:, :, sz["coef_single"]:]).view("<u2").reshape(-1)
        fl = read_floats_fp16(w, cesi.neig * nblk * nf * 2, cesi.FP16_COEF_EXP_SHARE_FLOATS)
        coefs[:, :, ns:] = fl.view(np.complex64).reshape(cesi.neig, nblk, nf).transpose(1, 0, 2)
    return bases, coefs


def load_compressed_eigen_vectors(path):
    """load_compressed_eigen_vectors (commpressed-eigen-io.cpp:1005-1055) for
    the whole lattice: (eigen_values, cesi, bases, coefs) with blocks in
    global order, or None when path/metadata.txt does not exist.  A file whose
    CRC-32 differs from its metadata entry raises ValueError (the reference
    asserts)."""
    if not os.path.exists(os.path.join(path, "metadata.txt")):
        return None
    vals = read_eigen_values(path)
    cesi = read_compressed_eigen_system_info(path)
    n_files = int(np.prod(cesi.total_node))
    n_block = int(np.prod(cesi.total_block))
    bases = np.empty((n_block, cesi.nkeep, cesi.basis_c_size), np.complex64)
    coefs = np.empty((n_block, cesi.neig, cesi.nkeep), np.complex64)
    for idx in range(n_files):
        fn = file_name(path, idx, n_files)
        with open(fn, "rb") as f:
            raw = f.read()
        if len(raw) != cesi.file_size():
            raise ValueError(f"load_compressed_eigen_vectors: '{fn}' holds {len(raw)} bytes, "
                             f"expected {cesi.file_size()}")
        crc = zlib.crc32(raw) & 0xFFFFFFFF
        if crc != cesi.crcs[idx]:
            raise ValueError(f"load_compressed_eigen_vectors: file-idx={idx} loaded={crc:08X} "
                             f"metadata={cesi.crcs[idx]:08X}")
        g = _node_block_indices(cesi, idx)
        bases[g], coefs[g] = _decode_file(cesi, raw)
    return vals, cesi, bases, coefs


def load_low_modes(path):
    """load_low_modes (dslash.cpp:9-23): the LowModes stored under path, or
    None for '' or '/dev/null' or a directory without metadata.txt."""
    if path in ("", "/dev/null"):
        return None
    r = load_compressed_eigen_vectors(path)
    if r is None:
        return None
    vals, cesi, bases, coefs = r
    lm = LowModes(cesi.block_site, vals, bases, coefs)
    lm.cesi = cesi
    return lm


def save_compressed_eigen_vectors(path, lm, total_site, ls, node_site=None, nkeep_single=None,
                                  fp16_coef_exp_share_floats=2):
    """Writes lm in the layout load_compressed_eigen_vectors reads (the
    reference writes it with save_node_data and
    write_compressed_eigen_system_info, commpressed-eigen-io.cpp:161-197,
    :714-746): FP32 for the first nkeep_single basis vectors and
    coefficients (all of them by default), 16-bit codes for the rest.
    Returns the CompressedEigenSystemInfo written."""
    total_site = tuple(int(v) for v in total_site)
    node_site = tuple(int(v) for v in (node_site or total_site))
    nkeep = lm.n_basis
    nkeep_single = nkeep if nkeep_single is None else int(nkeep_single)
    cesi = CompressedEigenSystemInfo(total_site, node_site, lm.block_site, ls, lm.n_vec, nkeep, nkeep_single,
                                     fp16_coef_exp_share_floats)
    if lm.bases.shape != (int(np.prod(cesi.total_block)), nkeep, cesi.basis_c_size):
        raise ValueError(f"save_compressed_eigen_vectors: bases of shape {lm.bases.shape} do not fit "
                         f"{total_site} Ls {ls} in blocks {lm.block_site}")
    os.makedirs(path, exist_ok=True)
    n_files = int(np.prod(cesi.total_node))
    ns, bve = nkeep_single, cesi.block_vol_eo
    crcs = []
    for idx in range(n_files):
        g = _node_block_indices(cesi, idx)
        nblk = len(g)
        b = lm.bases[g].reshape(nblk, nkeep, bve, ls, 12).transpose(0, 1, 3, 2, 4)  # s slowest in a basis
        c = lm.coefs[g].transpose(1, 0, 2)  # (neig, node blocks, nkeep)
        parts = [np.ascontiguousarray(b[:, :ns]).astype("<c8").tobytes()]
        if ns < nkeep:
            fl = np.ascontiguousarray(b[:, ns:]).view(np.float32)
            parts.append(write_floats_fp16(fl, 24).astype("<u2").tobytes())
        cs = np.ascontiguousarray(c[:, :, :ns]).astype("<c8").view(np.uint8).reshape(lm.n_vec, nblk, ns * 8)
        if ns < nkeep:
            fl = np.ascontiguousarray(c[:, :, ns:]).view(np.float32)
            cf = write_floats_fp16(fl, cesi.FP16_COEF_EXP_SHARE_FLOATS).astype("<u2")
            cf = cf.view(np.uint8).reshape(lm.n_vec, nblk, cesi.sizes()["coef_fp16"])
            cs = np.concatenate([cs, cf], axis=2)
        parts.append(cs.tobytes())
        raw = b"".join(parts)
        assert len(raw) == cesi.file_size()
        fn = file_name(path, idx, n_files)
        os.makedirs(os.path.dirname(fn), exist_ok=True)
        with open(fn, "wb") as f:
            f.write(raw)
        crcs.append(zlib.crc32(raw) & 0xFFFFFFFF)
    cesi.crcs = crcs
    write_eigen_values(path, lm.eigen_values)
    write_compressed_eigen_system_info(cesi, path)
    return cesi
