"""Reading matrices from files and writing results (the reference's src/io, include/hsolve/io.hpp).

load_matrix restates hs::load_matrix (include/hsolve/io.hpp:101-151, src/io/load.cpp):
  * formats by extension: .npz (scipy.sparse.save_npz: csr / csc / coo, matrix or array API,
    row/col or coords keys; bsr / dia are Unsupported), .npy (dense 2-D), .mtx (MatrixMarket
    coordinate real / double / integer / complex / pattern x general / symmetric / hermitian, or
    array format; skew-symmetric is InvalidInput);
  * CSR built with rows sorted and duplicates summed (note "duplicates"), exact zeros dropped
    (note "zeros"; dense inputs drop them silently: there they mark structural absence);
  * non-finite values, non-square shapes and orders above 2^31 are InvalidInput;
  * complex input whose imaginary parts are all 0 is returned as real (note "real");
  * one stored triangle is restored (note "storage"): detected when one strict triangle is empty
    and the other is not (storage "auto"), declared by the storage option (entries in the other
    strict triangle are InvalidInput) or by a MatrixMarket symmetric / hermitian banner; the
    other triangle is the conjugate transpose, the plain transpose for a "symmetric" banner (a
    complex symmetric file then fails the Hermitian check);
  * Hermitian check: defect ||A - A^H||_F / ||A||_F above hermitian_tol is NotHermitian; a
    positive defect replaces A by its Hermitian part (note "hermitian"); check_hermitian=False
    skips both (defect -1, note "unchecked").
Notes are "<tag>: <text>" in that order. npz archives are read by a small zip reader of this
module (validated end record, zip64, stored or deflated members, CRC-32 and exact sizes checked)
and every member by numpy's npy reader with allow_pickle=False (object arrays are refused); zip
and npy errors are HsolveIOError.
"""
from __future__ import annotations

import io as _stdio
import os
import zlib
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from .errors import HsolveIOError, InvalidInputError, NotHermitianError, UnsupportedError

MAX_ORDER = 2 ** 31
STORAGES = ("auto", "full", "upper", "lower")


@dataclass
class LoadedMatrix:
    """A matrix read from disk, restored to full Hermitian storage (io.hpp:84-100)."""
    matrix: sp.csr_matrix
    source_format: str                 # "npz-csr", "npz-csc", "npz-coo", "npy-dense", "mtx"
    storage: str = "full"              # triangle restored from: "upper", "lower" or "full"
    hermitian_defect: float = 0.0      # -1 when the check was skipped
    notes: list = field(default_factory=list)

    @property
    def is_complex(self) -> bool:
        return np.iscomplexobj(self.matrix.data)

    @property
    def order(self) -> int:
        return int(self.matrix.shape[0])


# ---------------------------------------------------------------------------------- readers

def _bytesio(data: bytes):
    return _stdio.BytesIO(data)


def read_npy(path) -> np.ndarray:
    """One .npy array (numpy's reader, no pickles)."""
    try:
        with open(path, "rb") as f:
            arr = np.lib.format.read_array(f, allow_pickle=False)
            if f.read(1):
                raise HsolveIOError(f"{path}: trailing bytes after the npy payload")
            return arr
    except HsolveIOError:
        raise
    except FileNotFoundError as e:
        raise HsolveIOError(f"{path}: cannot read ({e.strerror})") from None
    except ValueError as e:
        if "pickle" in str(e) or "Object arrays" in str(e):
            raise UnsupportedError(f"{path}: object arrays are not supported") from None
        raise HsolveIOError(f"{path}: malformed npy ({e})") from None
    except (OSError, EOFError) as e:
        raise HsolveIOError(f"{path}: malformed npy ({e})") from None


# zip records (APPNOTE 4.3): end of central directory, zip64 end record and locator, central and
# local file headers
_EOCD, _Z64_REC, _Z64_LOC = b"PK\x05\x06", b"PK\x06\x06", b"PK\x06\x07"
_CEN, _LOC = b"PK\x01\x02", b"PK\x03\x04"


def _u16(b, o):
    return int.from_bytes(b[o:o + 2], "little")


def _u32(b, o):
    return int.from_bytes(b[o:o + 4], "little")


def _u64(b, o):
    return int.from_bytes(b[o:o + 8], "little")


def _central_directory(buf: bytes, path):
    """(offset, size, entries) of the central directory. The end record is searched from the end
    of the file backwards; a candidate counts only if its directory lies right before it (or
    before its zip64 record) and starts with a central header, so an archive comment or appended
    bytes that happen to contain the signature do not mislead it."""
    n = len(buf)
    pos = buf.rfind(_EOCD)
    while pos >= 0:
        if pos + 22 <= n:
            disk, cd_disk = _u16(buf, pos + 4), _u16(buf, pos + 6)
            entries, cd_size, cd_off = _u16(buf, pos + 10), _u32(buf, pos + 12), _u32(buf, pos + 16)
            end = pos
            ok = True
            saturated = entries == 0xFFFF or cd_size == 0xFFFFFFFF or cd_off == 0xFFFFFFFF
            loc = pos - 20
            if loc >= 0 and buf[loc:loc + 4] == _Z64_LOC:
                # a zip64 end record (numpy writes one for every archive) holds the directory
                # figures and sits between the directory and this record
                rec = _u64(buf, loc + 8)
                if rec + 56 > n or buf[rec:rec + 4] != _Z64_REC:
                    ok = False
                else:
                    disk, cd_disk = _u32(buf, rec + 16), _u32(buf, rec + 20)
                    entries, cd_size, cd_off = _u64(buf, rec + 32), _u64(buf, rec + 40), _u64(buf, rec + 48)
                    end = rec
            elif saturated:
                ok = False
            if ok and cd_off + cd_size == end and (entries == 0 or buf[cd_off:cd_off + 4] == _CEN):
                if disk != 0 or cd_disk != 0:
                    raise UnsupportedError(f"{path}: multi-volume zip archives are not supported")
                return cd_off, cd_size, entries
        pos = buf.rfind(_EOCD, 0, pos)
    raise HsolveIOError(f"{path}: no zip end-of-central-directory record")


def _members(buf: bytes, path):
    """(name, method, flags, crc, csize, usize, local header offset) of every central entry."""
    cd_off, cd_size, entries = _central_directory(buf, path)
    out, p = [], cd_off
    for _ in range(entries):
        if buf[p:p + 4] != _CEN:
            raise HsolveIOError(f"{path}: malformed central directory")
        flags, method = _u16(buf, p + 8), _u16(buf, p + 10)
        crc, csize, usize = _u32(buf, p + 16), _u32(buf, p + 20), _u32(buf, p + 24)
        nlen, xlen, clen = _u16(buf, p + 28), _u16(buf, p + 30), _u16(buf, p + 32)
        disk, off = _u16(buf, p + 34), _u32(buf, p + 42)
        name = buf[p + 46:p + 46 + nlen].decode("utf-8", errors="replace")
        extra = buf[p + 46 + nlen:p + 46 + nlen + xlen]
        need = [usize == 0xFFFFFFFF, csize == 0xFFFFFFFF, off == 0xFFFFFFFF]
        if any(need):
            q, found = 0, False
            while q + 4 <= len(extra):
                tag, size = _u16(extra, q), _u16(extra, q + 2)
                if tag == 0x0001:
                    found, r = True, q + 4
                    if need[0]:
                        usize, r = _u64(extra, r), r + 8
                    if need[1]:
                        csize, r = _u64(extra, r), r + 8
                    if need[2]:
                        off = _u64(extra, r)
                    break
                q += 4 + size
            if not found:
                raise HsolveIOError(f"{path}: member {name} lacks its zip64 extra field")
        if disk not in (0, 0xFFFF):
            raise UnsupportedError(f"{path}: multi-volume zip archives are not supported")
        out.append((name, method, flags, crc, csize, usize, off))
        p += 46 + nlen + xlen + clen
    return out


def _member_bytes(buf: bytes, path, m) -> bytes:
    name, method, flags, crc, csize, usize, off = m
    if flags & 0x1:
        raise UnsupportedError(f"{path}: encrypted zip members are not supported")
    if method not in (0, 8):
        raise UnsupportedError(f"{path}: zip compression method {method} is not supported")
    if buf[off:off + 4] != _LOC:
        raise HsolveIOError(f"{path}: member {name}: bad local header")
    start = off + 30 + _u16(buf, off + 26) + _u16(buf, off + 28)
    comp = buf[start:start + csize]
    if len(comp) != csize:
        raise HsolveIOError(f"{path}: member {name} is truncated")
    if method == 0:
        data = comp
    else:
        d = zlib.decompressobj(-15)
        try:
            data = d.decompress(comp, usize + 1)
        except zlib.error as e:
            raise HsolveIOError(f"{path}: member {name}: bad deflate data ({e})") from None
        if len(data) > usize:
            raise HsolveIOError(f"{path}: member {name} inflates to more data than its declared size")
        if not d.eof:
            if len(data) == usize and not d.unconsumed_tail:
                raise HsolveIOError(f"{path}: member {name}: deflate data is truncated")
            raise HsolveIOError(f"{path}: member {name}: deflate data is truncated")
        if len(data) < usize:
            raise HsolveIOError(f"{path}: member {name}: deflate stream ends before the declared size")
    if len(data) != usize:
        raise HsolveIOError(f"{path}: member {name} has {len(data)} bytes, declared {usize}")
    if zlib.crc32(data) & 0xFFFFFFFF != crc:
        raise HsolveIOError(f"{path}: member {name}: CRC-32 mismatch")
    return data


def read_npz(path) -> dict:
    """Every numpy member of an .npz archive (keys without ".npy"; members whose names do not end in
    ".npy" are skipped, as np.load returns them as raw bytes). The archive is located by its
    validated end record (zip64 honoured), members are stored or deflated, CRC-32 and sizes are
    verified; each payload is parsed by numpy's npy reader without pickles."""
    try:
        with open(path, "rb") as f:
            buf = f.read()
    except FileNotFoundError as e:
        raise HsolveIOError(f"{path}: cannot read ({e.strerror})") from None
    except OSError as e:
        raise HsolveIOError(f"{path}: cannot read ({e})") from None
    members = _members(buf, path)
    names = [m[0] for m in members]
    if len(set(names)) != len(names):
        raise HsolveIOError(f"{path}: duplicate npz member")
    out = {}
    for m in members:
        if not m[0].endswith(".npy"):
            continue
        data = _member_bytes(buf, path, m)
        b = _bytesio(data)
        try:
            arr = np.lib.format.read_array(b, allow_pickle=False)
        except ValueError as e:
            if "pickle" in str(e) or "Object arrays" in str(e):
                raise UnsupportedError(f"{path}: object arrays are not supported") from None
            raise HsolveIOError(f"{path}: member {m[0]}: malformed npy ({e})") from None
        except (OSError, EOFError) as e:
            raise HsolveIOError(f"{path}: member {m[0]}: malformed npy ({e})") from None
        if b.read(1):
            raise HsolveIOError(f"{path}: trailing bytes in member {m[0]}")
        out[m[0][:-4]] = arr
    return out


def _value_array(a: np.ndarray, what: str) -> np.ndarray:
    """Values as float64 / complex128 (safe casts: bool, integers, f2/f4/f8, c8/c16)."""
    k = a.dtype.kind
    if k == "c":
        return a.astype(np.complex128)
    if k in "fiub":
        return a.astype(np.float64)
    raise InvalidInputError(f"{what}: unsupported value dtype {a.dtype.str}")


def _index_array(a: np.ndarray, what: str) -> np.ndarray:
    if a.dtype.kind not in "iu":
        raise InvalidInputError(f"{what}: index arrays must be integers, got {a.dtype.str}")
    return a.astype(np.int64)


class _Raw:
    """What a reader hands to the common pipeline: COO triplets of a square matrix."""

    def __init__(self, n, rows, cols, vals, fmt, declared="general", dense=False, notes=None):
        self.n, self.rows, self.cols, self.vals = n, rows, cols, vals
        self.format, self.declared, self.dense = fmt, declared, dense
        self.notes = notes or []


def _shape2(shape, what) -> int:
    shape = [int(x) for x in np.asarray(shape).ravel()]
    if len(shape) != 2:
        raise InvalidInputError(f"{what}: expected a 2-D matrix, got shape {tuple(shape)}")
    if shape[0] != shape[1]:
        raise InvalidInputError(f"{what}: matrix must be square, got shape {tuple(shape)}")
    if shape[0] < 1 or shape[0] > MAX_ORDER:
        raise InvalidInputError(f"{what}: order {shape[0]} is outside [1, 2^31]")
    return shape[0]


def _read_npz_matrix(path) -> _Raw:
    m = read_npz(path)
    if "format" not in m:
        raise InvalidInputError(f"{path}: not a scipy sparse archive (no 'format' member)")
    fmt = m["format"]
    fmt = fmt.item().decode() if fmt.dtype.kind == "S" else str(fmt.item())
    if fmt in ("bsr", "dia", "dok", "lil"):
        raise UnsupportedError(f"{path}: scipy '{fmt}' archives are not supported (save as csr)")
    if fmt not in ("csr", "csc", "coo"):
        raise InvalidInputError(f"{path}: unknown sparse format '{fmt}'")
    if "shape" not in m or "data" not in m:
        raise InvalidInputError(f"{path}: a '{fmt}' archive needs 'shape' and 'data' members")
    if np.asarray(m["shape"]).size != 2:
        raise InvalidInputError(f"{path}: expected a 2-D matrix, got shape {tuple(np.asarray(m['shape']).ravel())}")
    n = _shape2(m["shape"], path)
    vals = _value_array(np.asarray(m["data"]).ravel(), path)
    used = {"format", "shape", "data", "_is_array"}
    if fmt in ("csr", "csc"):
        for key in ("indices", "indptr"):
            if key not in m:
                raise InvalidInputError(f"{path}: a '{fmt}' archive needs an '{key}' member")
        ind = _index_array(m["indices"].ravel(), path)
        ptr = _index_array(m["indptr"].ravel(), path)
        used |= {"indices", "indptr"}
        if ptr.size != n + 1 or ptr[0] != 0 or np.any(np.diff(ptr) < 0) or ptr[-1] != ind.size or ind.size != vals.size:
            raise InvalidInputError(f"{path}: inconsistent '{fmt}' index arrays")
        if ind.size and (ind.min() < 0 or ind.max() >= n):
            raise InvalidInputError(f"{path}: index out of range")
        major = np.repeat(np.arange(n, dtype=np.int64), np.diff(ptr))
        rows, cols = (major, ind) if fmt == "csr" else (ind, major)
    else:
        if "coords" in m:
            c = _index_array(m["coords"], path)
            if c.ndim != 2 or c.shape[0] != 2:
                raise InvalidInputError(f"{path}: coo coords must be 2 x nnz")
            rows, cols = c[0], c[1]
            used.add("coords")
        elif "row" in m and "col" in m:
            rows, cols = _index_array(m["row"].ravel(), path), _index_array(m["col"].ravel(), path)
            used |= {"row", "col"}
        else:
            raise InvalidInputError(f"{path}: a 'coo' archive needs 'row'/'col' or 'coords'")
        if rows.size != vals.size or cols.size != vals.size:
            raise InvalidInputError(f"{path}: inconsistent 'coo' arrays")
        if rows.size and (min(rows.min(), cols.min()) < 0 or max(rows.max(), cols.max()) >= n):
            raise InvalidInputError(f"{path}: index out of range")
    notes = []
    extra = sorted(set(m) - used)
    if extra:
        notes.append("ignored: npz members " + ", ".join(extra) + " are not part of the matrix")
    return _Raw(n, rows, cols, vals, "npz-" + fmt, notes=notes)


def _read_npy_matrix(path) -> _Raw:
    a = read_npy(path)
    if a.ndim != 2:
        raise InvalidInputError(f"{path}: expected a 2-D array, got {a.ndim}-D")
    n = _shape2(a.shape, path)
    v = _value_array(a, path)
    r, c = np.nonzero(v)
    return _Raw(n, r.astype(np.int64), c.astype(np.int64), v[r, c], "npy-dense", dense=True)


def _read_mtx_matrix(path) -> _Raw:
    try:
        with open(path, "rb") as f:
            text = f.read().decode("ascii", errors="strict")
    except FileNotFoundError as e:
        raise HsolveIOError(f"{path}: cannot read ({e.strerror})") from None
    except (OSError, UnicodeDecodeError) as e:
        raise HsolveIOError(f"{path}: cannot read ({e})") from None
    lines = text.replace("\r\n", "\n").replace("\r", "\n").split("\n")
    if not lines or not lines[0].lower().startswith("%%matrixmarket"):
        raise HsolveIOError(f"{path}: missing %%MatrixMarket banner")
    head = lines[0].split()
    if len(head) != 5 or head[1].lower() != "matrix":
        raise HsolveIOError(f"{path}: malformed MatrixMarket banner")
    layout, field_, symm = head[2].lower(), head[3].lower(), head[4].lower()
    if layout not in ("coordinate", "array") or field_ not in ("real", "double", "integer", "complex", "pattern") \
            or symm not in ("general", "symmetric", "hermitian", "skew-symmetric"):
        raise HsolveIOError(f"{path}: unknown MatrixMarket type '{' '.join(head[2:])}'")
    if symm == "skew-symmetric":
        raise InvalidInputError(f"{path}: a skew-symmetric matrix is not Hermitian")
    if layout == "array" and field_ == "pattern":
        raise HsolveIOError(f"{path}: pattern field in array format")
    body = [ln for ln in lines[1:] if ln.strip() and not ln.lstrip().startswith("%")]
    if not body:
        raise HsolveIOError(f"{path}: missing size line")
    size = body[0].split()
    try:
        dims = [int(x) for x in size]
    except ValueError:
        raise HsolveIOError(f"{path}: malformed size line") from None
    ncomp = 2 if field_ == "complex" else (0 if field_ == "pattern" else 1)

    def number(tok):
        try:
            return float(tok)  # values below the double range read as 0, as scipy.io.mmread
        except ValueError:
            raise HsolveIOError(f"{path}: malformed value '{tok}'") from None

    entries = body[1:]
    if layout == "coordinate":
        if len(dims) != 3:
            raise HsolveIOError(f"{path}: coordinate size line needs rows cols entries")
        nr, nc, nnz = dims
        if nr != nc:
            raise InvalidInputError(f"{path}: matrix must be square, got shape ({nr}, {nc})")
        if len(entries) < nnz:
            raise HsolveIOError(f"{path}: truncated ({len(entries)} of {nnz} entries)")
        if len(entries) > nnz:
            raise HsolveIOError(f"{path}: more entries than declared")
        rows = np.empty(nnz, dtype=np.int64)
        cols = np.empty(nnz, dtype=np.int64)
        vals = np.empty(nnz, dtype=np.complex128 if ncomp == 2 else np.float64)
        for q, ln in enumerate(entries):
            tok = ln.split()
            if len(tok) != 2 + ncomp:
                raise HsolveIOError(f"{path}: malformed entry line {q + 1}")
            try:
                i, j = int(tok[0]), int(tok[1])
            except ValueError:
                raise HsolveIOError(f"{path}: malformed index on entry line {q + 1}") from None
            if not (1 <= i <= nr and 1 <= j <= nc):
                raise HsolveIOError(f"{path}: index ({i}, {j}) out of range on entry line {q + 1}")
            rows[q], cols[q] = i - 1, j - 1
            if ncomp == 0:
                vals[q] = 1.0
            elif ncomp == 1:
                vals[q] = number(tok[2])
            else:
                vals[q] = complex(number(tok[2]), number(tok[3]))
        n = nr
        if n < 1 or n > MAX_ORDER:
            raise InvalidInputError(f"{path}: order {n} is outside [1, 2^31]")
        dense = False
    else:
        if len(dims) != 2:
            raise HsolveIOError(f"{path}: array size line needs rows cols")
        nr, nc = dims
        if nr != nc:
            raise InvalidInputError(f"{path}: matrix must be square, got shape ({nr}, {nc})")
        n = nr
        # column-major; a symmetric / hermitian array stores the lower triangle incl. the diagonal
        pos = [(i, j) for j in range(n) for i in range(n if symm == "general" else n) if symm == "general" or i >= j]
        if len(entries) < len(pos):
            raise HsolveIOError(f"{path}: truncated ({len(entries)} of {len(pos)} values)")
        if len(entries) > len(pos):
            raise HsolveIOError(f"{path}: more values than the array holds")
        vals_all = []
        for q, ln in enumerate(entries):
            tok = ln.split()
            if len(tok) != ncomp:
                raise HsolveIOError(f"{path}: malformed value line {q + 1}")
            vals_all.append(number(tok[0]) if ncomp == 1 else complex(number(tok[0]), number(tok[1])))
        rr = np.array([p[0] for p in pos], dtype=np.int64)
        cc = np.array([p[1] for p in pos], dtype=np.int64)
        vv = np.array(vals_all, dtype=np.complex128 if ncomp == 2 else np.float64)
        keep = vv != 0
        rows, cols, vals = rr[keep], cc[keep], vv[keep]
        dense = True
    return _Raw(n, rows, cols, vals, "mtx", declared=symm, dense=dense)


# ---------------------------------------------------------------------------------- pipeline

def _strict_counts(M: sp.csr_matrix):
    coo = M.tocoo()
    return int(np.count_nonzero(coo.row > coo.col)), int(np.count_nonzero(coo.row < coo.col))


def _defect(M: sp.csr_matrix) -> float:
    fro = sp.linalg.norm(M) if M.nnz else 0.0
    D = (M - M.conj().T).tocsr()
    D.eliminate_zeros()
    if D.nnz == 0 or fro == 0.0:
        return 0.0
    return max(float(sp.linalg.norm(D) / fro), np.finfo(np.float64).tiny)


def load_matrix(path, storage: str = "auto", hermitian_tol: float = 1e-10, check_hermitian: bool = True,
                drop_zeros: bool = True) -> LoadedMatrix:
    """Read a square Hermitian matrix from .npz / .npy / .mtx (see the module docstring)."""
    path = os.fspath(path)
    if storage not in STORAGES:
        raise InvalidInputError(f"storage = {storage!r} is not one of {STORAGES}")
    ext = os.path.splitext(path)[1].lower()
    if ext == ".npz":
        raw = _read_npz_matrix(path)
    elif ext == ".npy":
        raw = _read_npy_matrix(path)
    elif ext == ".mtx":
        raw = _read_mtx_matrix(path)
    else:
        raise InvalidInputError(f"{path}: unknown matrix file extension (expected .npz, .npy or .mtx)")
    notes = list(raw.notes)
    n = raw.n
    vals = raw.vals
    if vals.size and not np.all(np.isfinite(vals)):
        raise InvalidInputError(f"{path}: non-finite value")
    # CSR: sum duplicates, drop exact zeros
    key = raw.rows * n + raw.cols
    uniq, inv = np.unique(key, return_inverse=True)
    dups = int(key.size - uniq.size)
    summed = np.zeros(uniq.size, dtype=vals.dtype)
    np.add.at(summed, inv, vals)
    if dups:
        notes.append(f"duplicates: summed {dups} duplicate entries")
    rows, cols = uniq // n, uniq % n
    if raw.dense or drop_zeros:
        nz = summed != 0
        zeros = int(summed.size - np.count_nonzero(nz))
        if zeros and not raw.dense:
            notes.append(f"zeros: dropped {zeros} explicit zero entries")
        rows, cols, summed = rows[nz], cols[nz], summed[nz]
    if np.iscomplexobj(summed) and not np.any(summed.imag != 0):
        summed = summed.real.copy()
        notes.append("real: complex input has all-zero imaginary parts; returned as real")
    M = sp.csr_matrix((summed, (rows, cols)), shape=(n, n))
    M.sort_indices()
    # storage
    lower, upper = _strict_counts(M)
    detected = "full"
    if raw.declared in ("symmetric", "hermitian"):
        if storage == "full":
            raise InvalidInputError(f"{path}: a {raw.declared} MatrixMarket file stores one triangle, not 'full'")
        if lower and upper:
            raise InvalidInputError(f"{path}: a {raw.declared} MatrixMarket file holds entries in both triangles")
        tri = "upper" if upper else "lower"
        if storage in ("upper", "lower") and (lower if storage == "upper" else upper):
            raise InvalidInputError(f"{path}: storage '{storage}' but entries in the other triangle")
        detected = storage if storage in ("upper", "lower") else tri
    elif storage in ("upper", "lower"):
        if (lower if storage == "upper" else upper):
            raise InvalidInputError(f"{path}: storage '{storage}' but entries in the other strict triangle")
        detected = storage
    elif storage == "auto":
        if upper and not lower:
            detected = "upper"
        elif lower and not upper:
            detected = "lower"
    if detected != "full":
        T = sp.triu(M) if detected == "upper" else sp.tril(M)
        strict = sp.triu(T, 1) if detected == "upper" else sp.tril(T, -1)
        # a MatrixMarket "symmetric" banner implies a_ji = a_ij, every other declaration a_ji = conj(a_ij)
        mirror = strict.T if raw.declared == "symmetric" else strict.conj().T
        M = (T + mirror).tocsr()
        M.sort_indices()
        notes.append(f"storage: {detected} triangle mirrored by Hermitian symmetry")
    # Hermitian check
    if not check_hermitian:
        notes.append("unchecked: Hermitian check skipped")
        return LoadedMatrix(M, raw.format, detected, -1.0, notes)
    d = _defect(M)
    if not (d <= hermitian_tol):
        raise NotHermitianError(f"{path}: not Hermitian: ||M - M^H||_F / ||M||_F = {d:.3g} > {hermitian_tol:.3g}")
    if d > 0.0:
        M = ((M + M.conj().T) * 0.5).tocsr()
        M.eliminate_zeros()
        M.sort_indices()
        notes.append(f"hermitian: Hermitian part taken (defect {d:.3g})")
    return LoadedMatrix(M, raw.format, detected, d, notes)


# ---------------------------------------------------------------------------------- writers

def write_npy_vector(path, v) -> None:
    """1-D array as np.save writes it (little-endian descr)."""
    np.save(path, np.ascontiguousarray(v))


def write_npy_matrix(path, m) -> None:
    """(rows, cols) matrix stored column-major, as np.save(np.asfortranarray(m)) writes it."""
    np.save(path, np.asfortranarray(m))
