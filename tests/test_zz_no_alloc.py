"""rds_cluster_labels_static -- the allocation-free variant of
rds_cluster_labels -- against the allocating entry point, through the C ABI.

Two calls in a row on one GPU, the second with a smaller system, each equal
to rds_cluster_labels on the same state (labels are integers: bit-exact); a
context above RDS_STATIC_LABELS_MAX particles is refused before any device
work.
"""

import ctypes as C

import numpy as np
import pytest
import torch

from helpers import engine_for, small
from pyrdsim_amd import _lib as L

pytestmark = pytest.mark.gpu


def _labels(lib, fn, eng, r_link, mask):
    out = torch.full((eng.n_total,), -7, dtype=torch.int32, device=eng.device)
    rc = getattr(lib, fn)(eng.ctx, C.c_double(r_link), C.c_uint32(mask), out.data_ptr())
    assert rc == L.OK, lib.rds_last_error().decode()
    torch.cuda.synchronize(eng.device)
    return out.cpu().numpy()


def test_static_labels_equal_allocating_entry_point(gpu):
    lib = L.load()
    big = small('wall', n=8192, seed=41)
    little = small('wall', n=2048, seed=42)
    for s, frac in ((big, 0.5), (little, 0.3)):
        rng = np.random.Generator(np.random.PCG64(len(s['x'])))
        s['types'] = (rng.random(len(s['x'])) < frac).astype(np.intc)
        eng = engine_for(s, precision='mixed')
        eng.run(5)
        eng.sync()
        for r_link, mask in ((1.5, 1 << 1), (1.2, 0)):
            want = _labels(lib, 'rds_cluster_labels', eng, r_link, mask)
            got = _labels(lib, 'rds_cluster_labels_static', eng, r_link, mask)
            assert np.array_equal(got, want), (len(s['x']), r_link, mask)
            # non-trivial clusters: some label is shared by several particles
            lab = want[want >= 0]
            assert len(lab) and np.bincount(lab).max() > 1
        eng.close()


def test_static_labels_refuse_more_than_the_maximum(gpu):
    from pyrdsim_amd.engine import Engine, build_model, make_replica
    lib = L.load()
    n = L.STATIC_LABELS_MAX + 1
    side = int(np.ceil(np.sqrt(n)))
    Lbox = side / np.sqrt(0.5)
    m = small('lj1')['model']
    model, rates = build_model(m['interactions'], m['reactions'], 1)
    rep = make_replica(0, n, [Lbox, Lbox], [1, 1], 1e-4, m['tau_by_tauD'], rates, 1)
    eng = Engine([model], [rep], precision='mixed')
    out = torch.empty((1,), dtype=torch.int32, device=eng.device)
    rc = lib.rds_cluster_labels_static(eng.ctx, C.c_double(1.5), C.c_uint32(0), out.data_ptr())
    assert rc == L.E_ARG
    assert 'at most %d particles' % L.STATIC_LABELS_MAX in lib.rds_last_error().decode()
    eng.close()
