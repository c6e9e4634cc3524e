"""Writes tests/golden/kitchen_mesh.npz from the crate's kitchen asset (assets/kitchen.obj.zst), the input of the
known-answer render test (examples/obj_cwbvh.rs:142-182, hash 1343358762).

The asset is a zstd-compressed Wavefront OBJ (Blender export).  It is decoded with pyarrow's zstd stream (the
crate's example uses ruzstd; both implement the zstd format, RFC 8878) and read the way
examples/helpers/load_obj.rs reads it with the `obj` crate: vertex positions from `v` lines parsed as f32 (correctly
rounded, like Rust's `str::parse::<f32>`; glibc strtof here), and for every face in file order (objects and groups
are concatenated in order by the example) the triangle (p0, p1, p2) plus (p0, p2, p3) for a quad.  The fixture keeps
the positions (f32 [P, 3]) and the faces as 0-based position indices (u32 [T, 3]); no texture or normal data is
used by the example.

Run in a container that has the reference checkout: python tests/golden/make_kitchen.py"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ASSET = "/root/reference/assets/kitchen.obj.zst"
OUT = os.path.join(HERE, "kitchen_mesh.npz")


def main():
    import pyarrow as pa
    raw = open(ASSET, "rb").read()
    text = pa.CompressedInputStream(pa.BufferReader(raw), "zstd").read().decode("utf-8")
    libc = ctypes.CDLL("libc.so.6")
    libc.strtof.restype = ctypes.c_float
    libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    pos, faces = [], []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            pos.append([libc.strtof(p.encode(), None) for p in parts[1:4]])
        elif parts[0] == "f":
            idx = []
            for p in parts[1:]:
                k = int(p.split("/")[0])
                idx.append(k - 1 if k > 0 else len(pos) + k)  # OBJ indices are 1-based; negative = relative
            faces.append([idx[0], idx[1], idx[2]])
            if len(idx) == 4:
                faces.append([idx[0], idx[2], idx[3]])
    positions = np.asarray(pos, dtype=np.float32)
    tri_idx = np.asarray(faces, dtype=np.uint32)
    assert tri_idx.max() < len(positions)
    np.savez_compressed(OUT, positions=positions, faces=tri_idx)
    print(f"{OUT}: {len(positions)} positions, {len(tri_idx)} triangles")
    return 0


if __name__ == "__main__":
    sys.exit(main())
