"""Convert the reference's reftest BSON files and published iteration tables into repo fixtures.

Run once, in a container that has the reference checkout (read as data only):

    python tests/golden/make_fixtures.py /root/reference

Inputs
  test/reftest-files/*.bson   -- BSON.jl documents {X, H}; each value is
                                 {tag:"array", type:{name:[Core,Float64]}, size:[...], data:<raw LE f64>}
                                 written by test/test_1D.jl:42-47, test/test_2D.jl, test/test_3D_*.jl
  output/out_diff_{1,2,3}D_{lin,linstep,nonlin}.txt
                              -- rows "nx [ny nz] ittot nt" appended by the converging scripts
                                 (e.g. scripts/diff_2D_lin.jl:84-87)

Outputs (tests/golden/)
  reftests.json     -- per fixture: sampled X and H (shape, column-major order) as float.hex strings,
                       sha256 of the source file
  iterations.json   -- published total PT iteration counts per (script, resolution)

The BSON reader below is a minimal decoder for the subset of the BSON spec that BSON.jl emits for
these files (documents, arrays, strings, int32/int64, doubles, binary); it executes nothing from the
file.
"""
import hashlib
import json
import os
import struct
import sys

import numpy as np


def _cstring(buf, pos):
    end = buf.index(b"\x00", pos)
    return buf[pos:end].decode("utf-8"), end + 1


def _document(buf, pos):
    (length,) = struct.unpack_from("<i", buf, pos)
    end = pos + length - 1  # trailing 0x00
    pos += 4
    out = {}
    while pos < end:
        etype = buf[pos]
        pos += 1
        key, pos = _cstring(buf, pos)
        if etype == 0x01:  # double
            (val,) = struct.unpack_from("<d", buf, pos)
            pos += 8
        elif etype == 0x02:  # string
            (slen,) = struct.unpack_from("<i", buf, pos)
            val = buf[pos + 4:pos + 4 + slen - 1].decode("utf-8")
            pos += 4 + slen
        elif etype in (0x03, 0x04):  # embedded document / array
            (dlen,) = struct.unpack_from("<i", buf, pos)
            val = _document(buf, pos)
            if etype == 0x04:
                val = [val[k] for k in sorted(val, key=int)]
            pos += dlen
        elif etype == 0x05:  # binary
            (blen,) = struct.unpack_from("<i", buf, pos)
            val = bytes(buf[pos + 5:pos + 5 + blen])
            pos += 5 + blen
        elif etype == 0x08:
            val = bool(buf[pos])
            pos += 1
        elif etype == 0x10:
            (val,) = struct.unpack_from("<i", buf, pos)
            pos += 4
        elif etype == 0x12:
            (val,) = struct.unpack_from("<q", buf, pos)
            pos += 8
        elif etype == 0x0A:
            val = None
        else:
            raise ValueError(f"unsupported BSON element type 0x{etype:02x} at {pos}")
        out[key] = val
    return out


def _julia_array(node):
    if node.get("tag") != "array":
        raise ValueError("expected a BSON.jl array node")
    name = node["type"]["name"]
    if name != ["Core", "Float64"]:
        raise ValueError(f"unexpected element type {name}")
    shape = tuple(int(s) for s in node["size"])
    data = np.frombuffer(node["data"], dtype="<f8")
    return data.reshape(shape, order="F")


def convert_reftests(ref_root):
    folder = os.path.join(ref_root, "test", "reftest-files")
    out = {}
    for fname in sorted(os.listdir(folder)):
        if not fname.endswith(".bson"):
            continue
        raw = open(os.path.join(folder, fname), "rb").read()
        doc = _document(raw, 0)
        entry = {"source": f"test/reftest-files/{fname}", "sha256": hashlib.sha256(raw).hexdigest()}
        for key in ("X", "H"):
            arr = _julia_array(doc[key])
            entry[key] = {
                "shape": list(arr.shape),
                "order": "F",
                "hex": [float(v).hex() for v in arr.ravel(order="F")],
            }
        out[fname[:-5]] = entry
    return out


def convert_iterations(ref_root):
    out = {}
    folder = os.path.join(ref_root, "output")
    for nd in (1, 2, 3):
        for case in ("lin", "linstep", "nonlin"):
            fname = f"out_diff_{nd}D_{case}.txt"
            rows = []
            for line in open(os.path.join(folder, fname)):
                vals = [int(v) for v in line.split()]
                if not vals:
                    continue
                res = vals[:nd]
                rows.append({"n": res, "ittot": vals[nd], "nt": vals[nd + 1]})
            out[f"{nd}D_{case}"] = {"source": f"output/{fname}", "rows": rows}
    return out


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "reftests.json"), "w") as f:
        json.dump(convert_reftests(ref_root), f, indent=1)
    with open(os.path.join(here, "iterations.json"), "w") as f:
        json.dump(convert_iterations(ref_root), f, indent=1)
    print("wrote reftests.json and iterations.json")


if __name__ == "__main__":
    main()
