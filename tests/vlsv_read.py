"""A VLSV reader for the tests, written from the format as the reference's
reader takes it (tools/rhybridpy.py:113-126, 263-290): the footer offset at
byte 8, an XML footer whose elements give each array's byte offset (text),
arraysize, vectorsize, datatype and datasize."""
import struct
import xml.etree.ElementTree as ET

import numpy as np

DT = {("float", 8): "<f8", ("float", 4): "<f4", ("uint", 8): "<u8", ("uint", 4): "<u4", ("uint", 1): "<u1",
      ("int", 8): "<i8", ("int", 4): "<i4"}


def read_vlsv(path):
    """{(tag, name): (array, attributes)} of every element of the footer."""
    raw = open(path, "rb").read()
    (off,) = struct.unpack("=Q", raw[8:16])
    root = ET.fromstring(raw[off:])
    out = {}
    for el in root:
        a = el.attrib
        n, v = int(a["arraysize"]), int(a["vectorsize"])
        data = np.frombuffer(raw, dtype=DT[(a["datatype"], int(a["datasize"]))], count=n * v, offset=int(el.text))
        if v > 1:
            data = data.reshape(n, v)
        out[(el.tag, a.get("name", ""))] = (data, a)
    return out
