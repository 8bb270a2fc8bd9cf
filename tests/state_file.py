"""Reader and writer of the library's binary state files (the layout that
adg_state_write / adg_state_read use), for golden fixtures written from the
oracle.  Little-endian: magic "ADGSTAT1"; int32 dim, k, m, nmax, ninst, 0;
int64 count[ninst]; per element int32 level[dim], int32 suppt[dim]; float64
coefficients, (k+1)^dim per element."""
import numpy as np

MAGIC = b"ADGSTAT1"


def pack(dim, k, m, nmax, instances):
    """instances: list of (level [n, dim], suppt [n, dim], ucoe [n * (k+1)^dim]) in hash-key order."""
    out = [MAGIC, np.array([dim, k, m, nmax, len(instances), 0], dtype="<i4").tobytes(),
           np.array([len(lv) for lv, _, _ in instances], dtype="<i8").tobytes()]
    for lv, sp, _ in instances:
        ls = np.stack([np.asarray(lv, dtype="<i4"), np.asarray(sp, dtype="<i4")], axis=1)  # [n, 2, dim]
        out.append(ls.tobytes())
    for _, _, u in instances:
        out.append(np.asarray(u, dtype="<f8").tobytes())
    return b"".join(out)


def unpack(data):
    assert data[:8] == MAGIC, "bad magic"
    dim, k, m, nmax, ninst, _ = np.frombuffer(data, dtype="<i4", count=6, offset=8)
    off = 8 + 24
    cnt = np.frombuffer(data, dtype="<i8", count=ninst, offset=off)
    off += 8 * int(ninst)
    K = (k + 1) ** dim
    insts = []
    for n in cnt:
        ls = np.frombuffer(data, dtype="<i4", count=int(n) * 2 * dim, offset=off).reshape(int(n), 2, dim)
        off += ls.nbytes
        insts.append([ls[:, 0, :].copy(), ls[:, 1, :].copy(), None])
    for i, n in enumerate(cnt):
        u = np.frombuffer(data, dtype="<f8", count=int(n) * K, offset=off)
        off += u.nbytes
        insts[i][2] = u.copy()
    assert off == len(data), "trailing bytes"
    return (int(dim), int(k), int(m), int(nmax)), [tuple(x) for x in insts]


def from_oracle(o):
    lv, sp, _ = o.grid()
    return pack(o.dim, o.k, o.m, o.nmax, [(lv, sp, o.get(0))])
