"""Core dump and load (api.save_cores / load_cores): one .npz of plain arrays per decomposition, the
engine's core layouts, readable with numpy.load(allow_pickle=False)."""
import numpy as np
import pytest

from oracle import ttsvd
from tensor_train_ring_decompositions_hpc_gpu_amd.api import load_cores, save_cores


def _random_tt(dims, ranks, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(ranks[k] * dims[k] * ranks[k + 1]).astype(np.float32) for k in range(len(dims))]


@pytest.mark.parametrize("tr", [False, True])
def test_round_trip_keeps_cores_and_tensor(tmp_path, tr):
    dims = [5, 4, 3, 6]
    ranks = [2, 3, 4, 2, 2] if tr else [1, 3, 4, 2, 1]
    cores = _random_tt(dims, ranks, 3)
    f = tmp_path / "cores.npz"
    # longer buffers (core_buffers hold the ceiling ranks) are cut to the ranks' sizes
    save_cores(f, [np.concatenate([c, np.ones(7, np.float32)]) for c in cores], dims, ranks, tr=tr)
    got, d2, r2, tr2 = load_cores(f, device=None)
    assert d2 == dims and r2 == ranks and tr2 == tr
    for a, b in zip(got, cores):
        assert a.dtype == np.float32 and np.array_equal(a, b)
    full = ttsvd.tr_full if tr else ttsvd.tt_full
    assert np.array_equal(full(got, dims, ranks), full(cores, dims, ranks))
    with np.load(f, allow_pickle=False) as z:      # nothing pickled
        assert all(z[k].dtype != object for k in z.files)


def test_rejects_short_cores(tmp_path):
    dims, ranks = [4, 4, 4], [1, 2, 2, 1]
    cores = _random_tt(dims, ranks, 1)
    cores[1] = cores[1][:-1]
    with pytest.raises(ValueError):
        save_cores(tmp_path / "x.npz", cores, dims, ranks)


@pytest.mark.gpu
def test_dump_of_a_decomposition_reconstructs(engine, tmp_path):
    dims, ranks = [12, 10, 8, 6], [1, 4, 5, 3, 1]
    A = engine.generate(dims, ranks, std=1.0, noise_rel=1e-6, seed=2)
    cores, sig, rep = engine.tt_svd(A, dims, ranks)
    f = tmp_path / "tt.npz"
    save_cores(f, cores, dims, rep["ranks"])
    got, d2, r2, tr = load_cores(f)
    e = engine.rel_error(got, d2, r2, A)
    assert not tr and e == pytest.approx(engine.rel_error(cores, dims, rep["ranks"], A), abs=1e-9) and e < 1e-4
