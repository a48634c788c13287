"""Readers and a formatter for the reference's output files (test
infrastructure).

es(v, w, d) restates the Fortran ESw.d edit descriptor on top of Python's
correctly rounded decimal conversion ('%.*E', round half to even): one digit
before the point, exponent E+zz for |e| <= 99 and +zzz (letter dropped) for
99 < |e| <= 999, right-justified in w characters, '*' * w when it does not
fit.  This is the number format of write_data_from_master
(io_nohdf5_m.f90:43 ASCII_FMT = 'ES14.6', 316-320) and of the header's
SOLUTIONTIME (ES18.4, io_nohdf5_m.f90:236).

read_tecplot parses the FTLE file of write_ftle (io_nohdf5_m.f90:183-265);
read_unformatted parses a Fortran unformatted sequential file with 4-byte
record markers (structured_io with TMP_FILE_FORMAT = BINARY,
io_nohdf5_m.f90:416-446,507-509).
"""
import math
import struct

import numpy as np


def es(v, w=14, d=6):
    v = float(v)
    if math.isnan(v):
        s = "NaN"
    elif math.isinf(v):
        s = "-Infinity" if v < 0 else "Infinity"
    else:
        t = "%.*E" % (d, v)
        mant, ex = t.split("E")
        e = int(ex)
        if abs(e) <= 99:
            s = mant + "E" + ("-" if e < 0 else "+") + "%02d" % abs(e)
        elif abs(e) <= 999:
            s = mant + ("-" if e < 0 else "+") + "%03d" % abs(e)
        else:
            return "*" * w
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def read_tecplot(path):
    """Returns (header lines, variable names, {name: list of 14-char fields},
    data lines) of a BLOCK-packed ES14.6 file."""
    with open(path, "r") as f:
        lines = f.read().split("\n")
    assert lines[-1] == "", "file must end with a newline"
    lines = lines[:-1]
    assert lines[0].startswith('TITLE = "') and lines[1] == "VARIABLES ="
    names = []
    i = 2
    while lines[i].startswith('"'):
        names.append(lines[i].strip('"'))
        i += 1
    header = lines[:i + 5]
    assert lines[i] == "ZONE"
    ijk = lines[i + 3]
    dims = [int(ijk[4:14]), int(ijk[19:29]), int(ijk[34:44])]
    npt = dims[0] * dims[1] * dims[2]
    data = lines[i + 5:]
    per_var = (npt + 9) // 10
    assert len(data) == per_var * len(names), (len(data), per_var, len(names))
    out = {}
    for k, name in enumerate(names):
        block = data[k * per_var:(k + 1) * per_var]
        fields = []
        for ln, line in enumerate(block):
            want = 10 if ln < per_var - 1 else npt - 10 * (per_var - 1)
            assert len(line) == 14 * want, (name, ln, len(line))
            fields.extend(line[14 * q:14 * (q + 1)] for q in range(want))
        out[name] = fields
    return header, names, out, dims


def read_unformatted(path, dtype):
    """Records of a Fortran unformatted sequential file as numpy arrays
    (single-subrecord records only)."""
    recs = []
    with open(path, "rb") as f:
        raw = f.read()
    pos = 0
    while pos < len(raw):
        (head,) = struct.unpack_from("<i", raw, pos)
        assert head >= 0, "continued records are not expected at these sizes"
        body = raw[pos + 4:pos + 4 + head]
        (tail,) = struct.unpack_from("<i", raw, pos + 4 + head)
        assert tail == head, (head, tail)
        recs.append(np.frombuffer(body, dtype=dtype).copy())
        pos += 8 + head
    return recs
