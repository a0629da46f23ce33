"""Seeded synthetic workloads of the benchmark configurations.

The reference reads edgelist files that it does not ship (README.md:7-17,
test_update_rpma.cu:174-192); these generators stand in for them, all drawn
from the library's MT19937-64 generators, so a seed fixes every key.
Generation runs on host threads (the ctypes calls release the GIL).
"""
import concurrent.futures
import math
import os

import numpy as np

from . import gen_rmat, gen_rmat_samples, shuffle

RMAT = (0.57, 0.19, 0.19)  # a, b, c; d = 0.05

_stream_keys = {}


def rmat_stream(scale=20, edges=1 << 24, seed=2, shuffle_seed=12, batch=100000, del_seed=13):
    """The R-MAT update stream of the power-law workload.

    `edges` distinct R-MAT edges on 2^scale vertices (a, b, c = RMAT, seed
    `seed`, self loops and repeats rejected), shuffled with `shuffle_seed`.
    The first half is the bulk load; the second half is cut into insert
    batches of `batch`; then as many edges as were inserted, a uniform sample
    without replacement (seed `del_seed`) of all `edges` -- the live set once
    the inserts are done -- are cut into delete batches of `batch`.
    Returns (bulk, insert batches, delete batches) as uint64 arrays.  The
    generated keys are cached per (scale, edges, seed, shuffle_seed).
    """
    k = (scale, edges, seed, shuffle_seed)
    if k not in _stream_keys:
        a, b, c = RMAT
        _stream_keys[k] = shuffle(gen_rmat(scale, edges, a, b, c, seed), shuffle_seed)
    keys = _stream_keys[k]
    half = edges // 2
    bulk = keys[:half]
    rest = keys[half:]
    ins = [rest[i:i + batch] for i in range(0, rest.size, batch)]
    rng = np.random.default_rng(del_seed)
    dels = keys[rng.permutation(edges)[: rest.size]]
    dls = [dels[i:i + batch] for i in range(0, dels.size, batch)]
    return bulk, ins, dls


def _pow2_scale(x):
    """log2 of the power of two nearest to x (at least 2^1)."""
    return max(1, int(round(math.log2(max(x, 2.0)))))


class BatchedWorkload:
    """Many independent R-MAT graphs, each receiving one mixed batch per round.

    Sizes: edge counts log-uniform in [e_min, e_max] (seed `seed`), vertex
    count = edges/16 rounded to the nearest power of two, graph i drawn with
    seed `seed + i` (distinct edges, no self loops).  Round r gives graph i a
    batch of lognormal size (mu = ln(mean_batch), sigma), capped at
    `batch_cap`, split `ins_frac` inserts (fresh R-MAT samples of the same
    graph, seed 100003 + i, repeats possible) and deletes drawn uniformly
    from the edges inserted so far (the initial graph and earlier insert
    batches; a drawn edge deleted earlier is a no-op).
    """

    def __init__(self, n_graphs=64, e_min=1e4, e_max=1e7, rounds=100, mean_batch=5000.0, sigma=1.5,
                 batch_cap=1_000_000, ins_frac=0.7, seed=3, threads=None):
        rng = np.random.default_rng(seed)
        self.n = n_graphs
        self.rounds = rounds
        self.edges = [int(round(10 ** x)) for x in rng.uniform(math.log10(e_min), math.log10(e_max), n_graphs)]
        self.scales = [_pow2_scale(e / 16.0) for e in self.edges]
        self.V = [1 << s for s in self.scales]
        sizes = np.minimum(np.round(rng.lognormal(math.log(mean_batch), sigma, (rounds, n_graphs))), batch_cap)
        self.sizes = np.maximum(sizes, 1).astype(np.int64)
        self.n_ins = np.round(self.sizes * ins_frac).astype(np.int64)
        self.n_del = self.sizes - self.n_ins
        self.seed = seed
        threads = threads or min(16, os.cpu_count() or 1)
        with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
            res = list(ex.map(self._graph, range(n_graphs)))
        self.initial = [r[0] for r in res]
        self.ins = [r[1] for r in res]   # ins[i][r]
        self.dels = [r[2] for r in res]  # dels[i][r]

    def _graph(self, i):
        s = self.scales[i]
        e = min(self.edges[i], (1 << s) * ((1 << s) - 1))
        a, b, c = RMAT
        initial = gen_rmat(s, e, a, b, c, self.seed + i)
        total_ins = int(self.n_ins[:, i].sum())
        stream = gen_rmat_samples(s, total_ins, a, b, c, 100003 + i)
        rng = np.random.default_rng(200003 + i)
        # the edges inserted so far are a prefix of initial ++ stream
        pool = np.concatenate([initial, stream])
        ins, dels = [], []
        off = 0
        for r in range(self.rounds):
            ni, nd = int(self.n_ins[r, i]), int(self.n_del[r, i])
            # deletes are drawn before this round's inserts join the pool
            dels.append(pool[rng.integers(0, initial.size + off, nd)])
            ins.append(stream[off:off + ni])
            off += ni
        return initial, ins, dels

    @property
    def total_updates(self):
        return int(self.sizes.sum())
