"""glibc's rand() and the reference harness's input cores, restated in numpy.

TEST INFRASTRUCTURE ONLY (the checker of the engine's reference-input generator,
``ttx_tt_generate_ref`` / ``ttx_tr_generate_ref``); only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg may import it.

* ``rand_stream(seed, count)`` -- glibc rand() after srand(seed): srandom_r's Park-Miller words
  r_0 = seed, r_i = 16807 r_{i-1} mod (2^31 - 1) (i < 31), then r_i = r_{i-31} (i = 31..33) and
  r_i = r_{i-3} + r_{i-31} mod 2^32 (random_r, TYPE_3); output k is r_{k+344} >> 1.  Pinned in the
  tests against the C library's own rand() through ctypes.
* ``uniform`` -- the reference's entry expression (float)(rand() / RAND_MAX) * 2.0 - 1.0, stored as
  float (generalTTTensor, ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:198-200).
* ``ref_cores`` -- generalTTTensor / generalTRTensor (ttdec6_singleGPU.cu:185-271,
  trdecomposition/high_order_GPUs/trdec6_singleGPU.cu:231-329): cores drawn in the order
  G_{d-1}, G_d, G_{d-2}, ..., G_1, each in its storage order.
"""
import numpy as np

RAND_MAX = 2147483647


def rand_stream(seed, count):
    """The first ``count`` values of glibc rand() after srand(seed) (uint32 array)."""
    w = int(seed) if seed != 0 else 1
    if w >= 2 ** 31:
        w -= 2 ** 32          # srand takes an unsigned int; srandom_r works on the signed word
    r = [w & 0xFFFFFFFF]
    for _ in range(1, 31):
        hi, lo = int(w / 127773), w - int(w / 127773) * 127773   # C division truncates towards zero
        w = 16807 * lo - 2836 * hi
        if w < 0:
            w += 2147483647
        r.append(w & 0xFFFFFFFF)
    total = 344 + count
    buf = np.zeros(total, dtype=np.uint64)
    buf[:31] = r
    for i in range(31, 34):
        buf[i] = buf[i - 31]
    for i in range(34, total):
        buf[i] = (buf[i - 3] + buf[i - 31]) & 0xFFFFFFFF
    return (buf[344:] >> np.uint64(1)).astype(np.uint32)


def uniform(vals):
    """(float)(v / RAND_MAX) * 2.0 - 1.0 rounded to float, as the reference stores it."""
    u = (np.asarray(vals, dtype=np.float64) / float(RAND_MAX)).astype(np.float32)
    return (u.astype(np.float64) * 2.0 - 1.0).astype(np.float32)


def core_sizes(dims, ranks, tr):
    d = len(dims)
    if not tr:
        return [ranks[k] * dims[k] * ranks[k + 1] for k in range(d)]
    out = [dims[0] * ranks[0] * ranks[1]]
    out += [ranks[k] * dims[k] * ranks[k + 1] for k in range(1, d - 1)]
    out += [ranks[d - 1] * dims[d - 1] * ranks[0]]
    return out


def draw_order(d):
    """Core indices (0-based) in the order the reference fills them: G_{d-1}, G_d, G_{d-2} .. G_1."""
    return [d - 2, d - 1] + list(range(d - 3, -1, -1))


def ref_cores(dims, ranks, tr=False, seed=1134):
    """Flat float32 cores in the engine's (= the reference's) storage layouts."""
    sizes = core_sizes(dims, ranks, tr)
    vals = uniform(rand_stream(seed, sum(sizes)))
    cores = [None] * len(dims)
    pos = 0
    for k in draw_order(len(dims)):
        cores[k] = vals[pos:pos + sizes[k]].copy()
        pos += sizes[k]
    return cores
