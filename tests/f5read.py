"""Reader of .f5 files (test infrastructure): the layout TACSFH5File writes
(src/io/TACSFH5.cpp:71-283) and scans (:340-420): an int count of
components, each as (int length with the terminating zero, name), then zones
of 5 ints (data type 0 int / 1 double / 2 float, dim1, dim2, zone-name
length, variable-names length), the two zero-terminated names and dim1 x dim2
values, native byte order."""
import numpy as np

_DT = {0: np.int32, 1: np.float64, 2: np.float32}


def read_f5(path):
    raw = open(path, "rb").read()
    off = 0

    def ints(n):
        nonlocal off
        v = np.frombuffer(raw, np.int32, n, off)
        off += 4 * n
        return v

    def text(n):
        nonlocal off
        s = raw[off:off + n]
        off += n
        assert s[-1:] == b"\0"
        return s[:-1].decode()

    ncomp = int(ints(1)[0])
    comps = [text(int(ints(1)[0])) for _ in range(ncomp)]
    zones = []
    while off < len(raw):
        dtype, dim1, dim2, lz, lv = (int(x) for x in ints(5))
        zname, vnames = text(lz), text(lv)
        dt = _DT[dtype]
        n = dim1 * dim2
        data = np.frombuffer(raw, dt, n, off).reshape(dim1, dim2)
        off += n * np.dtype(dt).itemsize
        zones.append({"name": zname, "vars": vnames.split(","), "type": dtype, "data": data})
    return comps, zones
