"""Read and write MRtrix .mif images from Python (the format of core/formats/mrtrix.cpp:40-117):
a text header starting "mrtrix image", "key: value" lines, "file: . <offset>", "END", then raw
data.  Supports Float32/Float64/(U)Int8/16/32 in either byte order and arbitrary layouts.

Arrays are returned in canonical numpy order data[x, y, z, (volume)] as float32.
"""
import numpy as np

_DT = {"float32": "f4", "float64": "f8", "int8": "i1", "uint8": "u1", "int16": "i2", "uint16": "u2",
       "int32": "i4", "uint32": "u4"}


def read(path):
    """Returns (data, header dict).  header keys: dim, vox, layout, transform, keyval (list)."""
    with open(path, "rb") as f:
        raw = f.read()
    end = raw.index(b"\nEND\n")
    lines = raw[:end].decode("latin-1").split("\n")
    if lines[0].strip() != "mrtrix image":
        raise ValueError("%s is not an MRtrix image" % path)
    h = dict(transform=[], keyval=[])
    for line in lines[1:]:
        if ":" not in line:
            continue
        k, v = line.split(":", 1)
        k, v = k.strip(), v.strip()
        lk = k.lower()
        if lk == "dim":
            h["dim"] = [int(x) for x in v.split(",")]
        elif lk == "vox":
            h["vox"] = [float(x) for x in v.split(",")]
        elif lk == "layout":
            h["layout"] = [x.strip() for x in v.split(",")]
        elif lk == "datatype":
            h["datatype"] = v
        elif lk == "transform":
            h["transform"].append([float(x) for x in v.split(",")])
        elif lk == "file":
            h["file"] = v
        elif lk == "scaling":
            h["scaling"] = [float(x) for x in v.split(",")]
        else:
            h["keyval"].append((k, v))
    dt = h["datatype"].lower()
    order = ">" if dt.endswith("be") else "<"
    base = dt[:-2] if dt[-2:] in ("le", "be") else dt
    fname, off = h["file"].split()[:2] if len(h["file"].split()) > 1 else (h["file"], "0")
    if fname != ".":
        import os
        with open(os.path.join(os.path.dirname(path), fname), "rb") as f:
            raw = f.read()
    dim = h["dim"]
    n = int(np.prod(dim))
    arr = np.frombuffer(raw, dtype=order + _DT[base], count=n, offset=int(off)).astype(np.float64)
    # file order: axes sorted by layout rank, rank 0 fastest
    ranks = [int(x[1:]) if x[0] in "+-" else int(x) for x in h["layout"]]
    signs = [-1 if x[0] == "-" else 1 for x in h["layout"]]
    axes = sorted(range(len(dim)), key=lambda a: ranks[a])          # fastest first
    arr = arr.reshape([dim[a] for a in reversed(axes)])             # C order: slowest first
    arr = np.transpose(arr, [len(dim) - 1 - axes.index(a) for a in range(len(dim))])
    for a, s in enumerate(signs):
        if s < 0:
            arr = np.flip(arr, axis=a)
    if "scaling" in h:
        arr = h["scaling"][0] + h["scaling"][1] * arr
    return np.ascontiguousarray(arr.astype(np.float32)), h


def write(path, data, vox=None, keyval=(), volume_contiguous=True, transform=None):
    """Write float32 data [x, y, z, (v)] with the volume axis contiguous (layout +1,+2,+3,+0)."""
    data = np.asarray(data, dtype=np.float32)
    nd = data.ndim
    vox = list(vox) if vox is not None else [1.0] * nd
    if volume_contiguous and nd == 4:
        layout = ["+1", "+2", "+3", "+0"]
        filed = np.transpose(data, (2, 1, 0, 3))                   # C order z, y, x, v
    else:
        layout = ["+%d" % a for a in range(nd)]
        filed = np.transpose(data, list(reversed(range(nd))))
    T = transform if transform is not None else [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]]
    head = "mrtrix image\ndim: %s\nvox: %s\nlayout: %s\ndatatype: Float32LE\n" % (
        ",".join(str(d) for d in data.shape), ",".join(repr(float(v)) for v in vox[:nd] + [1.0] * (nd - len(vox))),
        ",".join(layout))
    head += "".join("transform: %s\n" % ", ".join(repr(float(x)) for x in row) for row in T)
    for k, v in keyval:
        for line in str(v).split("\n"):
            head += "%s: %s\n" % (k, line)
    head += "file: "
    off = len(head.encode()) + 18
    off += (4 - off % 4) % 4
    head += ". %d\nEND\n" % off
    hb = head.encode()
    with open(path, "wb") as f:
        f.write(hb)
        f.write(b"\0" * (off - len(hb)))
        f.write(np.ascontiguousarray(filed).astype("<f4").tobytes())


def scheme_text(grad):
    return "\n".join(",".join(repr(float(x)) for x in row) for row in np.asarray(grad))
