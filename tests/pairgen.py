"""Synthetic pairs for the GPU tests at workload sizes, generated by a process pool.

The pool is started by conftest.py while the test session is collected, before any test
initialises HIP, so its forked workers hold no GPU device handles; without it (a module run on
its own) a pool is made per call, which is only safe while this process has not touched the
GPU either."""
from hbp_amd import synth

_pool = None


def start(workers=8):
    global _pool
    if _pool is None:
        import multiprocessing as mp
        _pool = mp.get_context("fork").Pool(workers)
    return _pool


def stop():
    global _pool
    if _pool is not None:
        _pool.close()
        _pool.join()
        _pool = None


def pair(seed, width, height, ndisp, n_planes=(6, 12)):
    """(left, right) of synth.make_pair, rows generated in parallel."""
    l, r, _ = synth.make_pair_parallel(seed, width, height, ndisp, n_planes, pool=_pool, workers=8)
    return l, r


def pairs(specs, n_planes=(6, 12)):
    return synth.make_pairs_parallel(specs, n_planes, pool=_pool, workers=8)
