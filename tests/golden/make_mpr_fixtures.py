"""Extract compact known-answer fixtures for the MPR contact stage from the
reference's recorded kernel launches.

Source: /root/reference/data/unittest_data.zip, files
mprPenetrationKernel{60,61,70,128}.bin -- serialized arguments of
mprPenetrationKernel (layout: b3LauncherCL::deserializeArgs,
src/Bullet3OpenCL/ParallelPrimitives/b3LauncherCL.cpp:114-152: an int argument
count, then per argument a 32-byte header {isBuffer, argIndex, sizeInBytes,
pad, 16 bytes of inline data} followed by the buffer bytes, then the trailing
global work size).  The reference test
test/OpenCL/AllBullet3Kernels/testExecuteBullet3NarrowphaseKernels.cpp:396-418
expects the contact counter (argument 8) to read 0, 1, 46 and 98 after the launch.

Only data is read (no code).  Each fixture keeps the pairs and the bodies,
collidables, hull records and hull vertices they reference, re-indexed densely.
Run here:  python tests/golden/make_mpr_fixtures.py
"""
import os
import struct
import zipfile

import numpy as np

ZIP = "/root/reference/data/unittest_data.zip"
OUT = os.path.dirname(os.path.abspath(__file__))
CASES = [("mprPenetrationKernel60.bin", 0), ("mprPenetrationKernel61.bin", 1),
         ("mprPenetrationKernel70.bin", 46), ("mprPenetrationKernel128.bin", 98)]


def load_args(buf):
    nargs = struct.unpack_from("<i", buf, 0)[0]
    idx = 4
    args = []
    for _ in range(nargs):
        is_buf, _argi, size, _pad = struct.unpack_from("<4i", buf, idx)
        inline = buf[idx + 16: idx + 32]
        idx += 32
        if is_buf:
            args.append(buf[idx: idx + size])
            idx += size
        else:
            args.append(inline[:size])
    return args


def main():
    z = zipfile.ZipFile(ZIP)
    for name, expected in CASES:
        a = load_args(z.read(name))
        assert len(a) == 11
        n_pairs = struct.unpack("<i", a[10])[0]
        pairs = np.frombuffer(a[0], np.int32).reshape(-1, 4)[:n_pairs].copy()
        bodies = np.frombuffer(a[1], np.uint8).reshape(-1, 80)
        cols = np.frombuffer(a[2], np.int32).reshape(-1, 4)
        convex = np.frombuffer(a[3], np.uint8).reshape(-1, 96)
        verts = np.frombuffer(a[4], np.float32).reshape(-1, 4)
        counter0 = struct.unpack("<i", a[8])[0]
        assert counter0 == 0

        used_b = np.unique(pairs[:, :2])
        bmap = {int(b): i for i, b in enumerate(used_b)}
        body_col = bodies[used_b].view(np.int32)[:, 16]
        used_c = np.unique(body_col)
        cmap = {int(c): i for i, c in enumerate(used_c)}
        col_shape = cols[used_c, 3]
        used_s = np.unique(col_shape)
        smap = {int(s): i for i, s in enumerate(used_s)}

        new_bodies = bodies[used_b].copy()
        nbi = new_bodies.view(np.int32)
        nbi[:, 16] = [cmap[int(c)] for c in body_col]
        new_cols = cols[used_c].copy()
        new_cols[:, 3] = [smap[int(s)] for s in col_shape]
        new_convex = convex[used_s].copy()
        nci = new_convex.view(np.int32)
        vlist = []
        off = 0
        for k, s in enumerate(used_s):
            ci = convex[s].view(np.int32)
            voff, nv = int(ci[20]), int(ci[19])
            vlist.append(verts[voff: voff + nv])
            nci[k, 20] = off
            nci[k, 17] = 0  # face offset (faces are not part of the recorded launch)
            nci[k, 18] = 0  # numFaces
            nci[k, 21] = 0  # unique edge offset
            nci[k, 22] = 0  # numUniqueEdges
            off += nv
        new_verts = np.concatenate(vlist).astype(np.float32)
        new_pairs = pairs.copy()
        new_pairs[:, 0] = [bmap[int(x)] for x in pairs[:, 0]]
        new_pairs[:, 1] = [bmap[int(x)] for x in pairs[:, 1]]
        new_pairs[:, 2:] = -1
        out = os.path.join(OUT, name.replace(".bin", ".npz"))
        np.savez_compressed(out, pairs=new_pairs, bodies=new_bodies, collidables=new_cols,
                            convexes=new_convex, vertices=new_verts,
                            expected_contacts=np.int32(expected),
                            orig_body_ids=used_b.astype(np.int32))
        print(name, "pairs", n_pairs, "bodies", len(used_b), "shapes", len(used_s),
              "verts", len(new_verts), "->", os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
