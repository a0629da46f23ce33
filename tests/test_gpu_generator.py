"""The device R-MAT generator against an independent restatement of its
definition (SplitMix64 counter stream, two levels per 64-bit draw), and its
quadrant statistics."""
import numpy as np
import pytest

import lpma_amd as L

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def rmat_reference(scale, n, a, b, c, seed):
    ta, tab, tabc = int(a * 2 ** 32), int((a + b) * 2 ** 32), int((a + b + c) * 2 ** 32)
    base = splitmix64(seed)
    out = []
    for i in range(n):
        for attempt in range(16):
            stream = splitmix64(base ^ ((i * 16 + attempt) & M64))
            src = dst = 0
            for l in range(0, scale, 2):
                r = splitmix64((stream + l) & M64)
                for h in range(2):
                    if l + h >= scale:
                        break
                    u = (r >> 32) if h else (r & 0xFFFFFFFF)
                    bs = 1 if u >= tab else 0
                    bd = 1 if (ta <= u < tab) or u >= tabc else 0
                    src = (src << 1) | bs
                    dst = (dst << 1) | bd
            if src != dst:
                break
        if src == dst:
            dst ^= 1
        out.append((src << 32) | dst)
    return np.array(out, dtype=np.uint64)


@pytest.mark.parametrize("scale,seed", [(3, 4), (10, 5), (24, 4)])
def test_device_rmat_matches_restatement(scale, seed):
    n = 3000
    got = L.to_host_u64(L.gen_rmat_device(scale, n, 0.57, 0.19, 0.19, seed))
    assert np.array_equal(got, rmat_reference(scale, n, 0.57, 0.19, 0.19, seed))


def test_device_rmat_statistics():
    scale, n = 16, 1 << 20
    k = L.to_host_u64(L.gen_rmat_device(scale, n, 0.57, 0.19, 0.19, 9))
    src, dst = k >> np.uint64(32), k & np.uint64(0xFFFFFFFF)
    assert (src < (1 << scale)).all() and (dst < (1 << scale)).all() and (src != dst).all()
    ts = (src >> np.uint64(scale - 1)) & np.uint64(1)
    td = (dst >> np.uint64(scale - 1)) & np.uint64(1)
    frac = [np.mean((ts == i) & (td == j)) for i in (0, 1) for j in (0, 1)]
    for f, p in zip(frac, (0.57, 0.19, 0.19, 0.05)):
        assert abs(f - p) < 0.01, frac
