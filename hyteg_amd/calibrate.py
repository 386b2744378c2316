"""Plain device copy of known size, used to calibrate rocprofv3 FETCH_SIZE /
WRITE_SIZE counters: `copy_from` between two FP64 functions on the cube at
the given level is one elementwise copy kernel (kCast<double,double>) over
nstore values (8*nstore bytes read, 8*nstore bytes written).

    python -m hyteg_amd.calibrate [--level 9] [--reps 5]
"""
import argparse
import json

import hyteg_amd as hg


def tet(n):
    return 0 if n <= 0 else n * (n + 1) * (n + 2) // 6


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--level", type=int, default=9)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    ctx = hg.Context(0)
    st = hg.PrimitiveStorage.cuboid(ctx)
    f = hg.P1Function("a", st, a.level, a.level)
    g = hg.P1Function("b", st, a.level, a.level)
    for _ in range(a.reps):
        g.copy_from(f, a.level)
    ctx.synchronize()
    m = 2 ** a.level
    nv, ne, nf = 8, 19, 18
    nstore = 6 * tet(m + 1) + nv + ne * (m - 1) + nf * ((m - 2) * (m - 1) // 2)
    print(json.dumps({"kernel": "kCast<double,double>", "elements": nstore, "bytes_read": 8 * nstore,
                      "bytes_written": 8 * nstore}))


if __name__ == "__main__":
    main()
