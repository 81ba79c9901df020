"""Matrix Market reading and writing with MTL4's semantics
(io/matrix_market.hpp:44-214 read, :218-300 write).

Header "%%MatrixMarket matrix <coordinate|array> <real|integer|pattern|complex>
<general|symmetric|skew-symmetric|hermitian>" (layout and symmetry compared
case-insensitively), '%' comment lines, then "rows cols [entries]".  Every
entry goes through an inserter: a later entry at the same position
overwrites (update_store); off-diagonal entries of symmetric and Hermitian
files are mirrored, of skew-symmetric files mirrored negated; pattern entries
are stored with value 0; complex values cannot be read into a real matrix;
array files are read column by column.  Result: CSR with sorted rows.
"""
import numpy as np

from ._lib import MtlError


class io_error(MtlError):
    pass


class file_not_found(io_error):
    pass


class runtime_error(MtlError):
    pass


def read_matrix_market(path):
    """Returns (nrows, ncols, rowptr int32, col int32, val float64)."""
    try:
        text = open(path).read()
    except OSError as e:
        raise file_not_found("matrix_market_istream: Error in input stream!\n"
                             "Probably file %s not found.\n" % path) from e
    tok = text.split(None, 5)
    if len(tok) < 5 or tok[0] != "%%MatrixMarket":
        raise runtime_error("File not in Matrix Market format")
    if tok[1] != "matrix":
        raise runtime_error("Try to read matrix from non-matrix file")
    layout, field, sym = tok[2].lower(), tok[3], tok[4].lower()
    if sym not in ("general", "symmetric", "skew-symmetric", "hermitian"):
        raise runtime_error("Unknown symmetry in Matrix Market file")
    if layout not in ("coordinate", "array"):
        raise runtime_error("Unknown sparsity in Matrix Market file")
    # body after the header line; skip comment lines
    body = text.split("\n", 1)[1] if "\n" in text else ""
    lines = [ln for ln in body.split("\n") if not ln.lstrip().startswith("%")]
    words = " ".join(lines).split()
    pos = 0
    nr, nc = int(words[0]), int(words[1])
    pos = 2
    if field == "real":
        kind = 0
    elif field == "integer":
        kind = 1
    elif field == "pattern":
        kind = 2
    elif field == "complex":
        raise runtime_error("Cannot convert complex value in real\n")
    else:
        raise runtime_error("Unknown tag for matrix value type in file")
    entries = {}

    def put(r, c, v):
        if not (0 <= r < nr and 0 <= c < nc):
            raise IndexError("Index out of range")
        entries[(r, c)] = v
        if r != c:
            if sym in ("symmetric", "hermitian"):
                entries[(c, r)] = v
            elif sym == "skew-symmetric":
                entries[(c, r)] = -v

    def value():
        nonlocal pos
        if kind == 2:
            return 0.0
        w = words[pos]
        pos += 1
        return float(int(w)) if kind == 1 else float(w)

    if layout == "coordinate":
        nnz = int(words[pos])
        pos += 1
        for _ in range(nnz):
            if pos + 1 >= len(words):
                break
            r, c = int(words[pos]) - 1, int(words[pos + 1]) - 1
            pos += 2
            put(r, c, value())
    else:
        for c in range(nc):
            for r in range(nr):
                put(r, c, value())
    keys = sorted(entries)
    rowptr = np.zeros(nr + 1, np.int64)
    for (r, _c) in keys:
        rowptr[r + 1] += 1
    rowptr = np.cumsum(rowptr).astype(np.int32)
    col = np.array([c for (_r, c) in keys], np.int32)
    val = np.array([entries[k] for k in keys], np.float64)
    return nr, nc, rowptr, col, val


def write_matrix_market(path, nrows, ncols, rowptr, col, val):
    """Coordinate real general with 17 significant digits (exact round trip)."""
    with open(path, "w") as f:
        f.write("%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n" % (nrows, ncols, len(col)))
        for r in range(nrows):
            for j in range(rowptr[r], rowptr[r + 1]):
                f.write("%d %d %s\n" % (r + 1, col[j] + 1, repr(float(val[j]))))
