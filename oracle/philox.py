"""Counter-based random numbers for the oracle: Philox4x32-10 and a bit-reproducible Box-Muller.

TEST INFRASTRUCTURE ONLY.  This module is the CPU checker for the engine's random sketch and
synthetic tensors; only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import it.  The shipped engine generates the same numbers with its own HIP code.

Why a project-owned generator: the reference draws its Gaussian sketch with cuRAND
(``curandGenerateNormal`` under ``CURAND_RNG_PSEUDO_DEFAULT``, seed 1234,
ttdecomposition/third_order_GPUs/opt/ttdec_gpu_float_rsvd_lu_half.cu:585-587) and its inputs with
glibc ``rand()`` after ``srand(1134)`` (ttdecomposition/high_order_GPUs/ttdec6_singleGPU.cu:986).
cuRAND's stream is not reproducible here, so the sketch Omega is defined by the published
Philox4x32-10 bijection (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
1, 2, 3", SC'11) and the Box-Muller transform (Box and Muller, 1958).  Every floating-point step
below is a single IEEE-754 binary32 operation (add, sub, mul, div, sqrt; no fused multiply-add),
so the HIP kernels reproduce the values bit for bit.

Definitions (shared with the HIP code):

* ``philox4x32_10(ctr[4], key[2])``: ten rounds; each round maps
  (c0,c1,c2,c3) -> (hi(M1*c2)^c1^k0, lo(M1*c2), hi(M0*c0)^c3^k1, lo(M0*c0)) and then bumps the
  key by (W0, W1).
* ``normal4(seed, stream, row, group)``: counter (row_lo, row_hi, group, stream), key
  (seed_lo, seed_hi); outputs x0..x3 -> Box-Muller pairs (x0,x1) and (x2,x3).
* Box-Muller pair (xa, xb): u = (2*(xa>>9)+1) * 2^-24 in (0,1); r = sqrt(-2*ln(u));
  angle = 2*pi*(xb>>8)/2^24 reduced to a quadrant; z0 = r*cos, z1 = r*sin.
* The sketch matrix Omega (n x ell) takes element (j, c) from normal4(seed, stream, j, c>>2)[c&3]
  rounded to bfloat16 (round to nearest even), with |value| < 2^-17 set to zero, so that every
  entry is exactly representable in bfloat16, float16 and float32.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)

# float32 constants; the HIP code uses the same literals rounded to binary32
F32 = np.float32
LN2 = F32(0.693147180559945309)
SQRT2_BITS = 0x3FB504F3            # mantissa threshold: f > sqrt(2) -> halve


def _q(num, den):
    """num/den as one binary32 division of exact binary32 operands (as the C literal 1.0f/11.0f)."""
    return F32(num) / F32(den)


LOG_C = [_q(1, 11), _q(1, 9), _q(1, 7), _q(1, 5), _q(1, 3)]
SIN_C = [_q(-1, 39916800), _q(1, 362880), _q(-1, 5040), _q(1, 120), _q(-1, 6)]
COS_C = [_q(1, 479001600), _q(-1, 3628800), _q(1, 40320), _q(-1, 720), _q(1, 24), F32(-0.5)]
ANG_SCALE = F32(2.0 * np.pi / 16777216.0)   # 2*pi / 2^24
U_SCALE = F32(1.0 / 16777216.0)             # 2^-24
BF16_FLUSH = F32(2.0 ** -17)


def philox4x32_10(ctr, key):
    """Philox4x32-10 on arrays. ctr: tuple of 4 uint32 arrays, key: tuple of 2 uint32 arrays."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint32) for c in ctr)
    k0, k1 = (np.asarray(k, dtype=np.uint32) for k in key)
    with np.errstate(over="ignore"):
        for rnd in range(10):
            p0 = c0.astype(np.uint64) * M0
            p1 = c2.astype(np.uint64) * M1
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & MASK32).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & MASK32).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            if rnd < 9:
                k0 = k0 + W0
                k1 = k1 + W1
    return c0, c1, c2, c3


def _as_f32(bits):
    return np.asarray(bits, dtype=np.uint32).view(np.float32)


def _bits(x):
    return np.asarray(x, dtype=np.float32).view(np.uint32)


def log_f32(u):
    """Natural log of u in (0, 1] as a fixed sequence of binary32 operations.

    u = 2^e * f with f in [sqrt(1/2), sqrt(2)); ln f = 2*atanh(s), s = (f-1)/(f+1), atanh by its
    odd Taylor series through s^11 (|s| <= 0.1716, truncation below 1e-11)."""
    u = np.asarray(u, dtype=np.float32)
    b = _bits(u)
    e = (b >> np.uint32(23)).astype(np.int32) - 127
    mbits = (b & np.uint32(0x7FFFFF)) | np.uint32(0x3F800000)
    big = mbits > np.uint32(SQRT2_BITS)
    mbits = np.where(big, mbits - np.uint32(0x00800000), mbits)   # f/2, exact
    e = np.where(big, e + 1, e)
    f = _as_f32(mbits)
    one = F32(1.0)
    s = (f - one) / (f + one)
    s2 = s * s
    p = LOG_C[0]
    for c in LOG_C[1:]:
        p = p * s2 + c              # two roundings: mul then add (no FMA)
    p = p * s2
    lnf = (s + s * p) * F32(2.0)
    return e.astype(np.float32) * LN2 + lnf


def sincos_quadrant(xb):
    """(cos, sin) of 2*pi*(xb>>8)/2^24 with exact quadrant reduction and Taylor polynomials."""
    w = (np.asarray(xb, dtype=np.uint32) >> np.uint32(8)).astype(np.int64)
    q = (w + (1 << 21)) >> 22                      # nearest quadrant, 0..4
    d = (w - (q << 22)).astype(np.float32)         # integer in [-2^21, 2^21), exact in binary32
    phi = d * ANG_SCALE
    p2 = phi * phi
    hs = SIN_C[0]
    for c in SIN_C[1:]:
        hs = hs * p2 + c
    sn = phi + (phi * p2) * hs
    hc = COS_C[0]
    for c in COS_C[1:]:
        hc = hc * p2 + c
    cs = F32(1.0) + p2 * hc
    q = q & 3
    cos_t = np.where(q == 0, cs, np.where(q == 1, -sn, np.where(q == 2, -cs, sn)))
    sin_t = np.where(q == 0, sn, np.where(q == 1, cs, np.where(q == 2, -sn, -cs)))
    return cos_t.astype(np.float32), sin_t.astype(np.float32)


def box_muller(xa, xb):
    t = (np.asarray(xa, dtype=np.uint32) >> np.uint32(9)).astype(np.uint32) * np.uint32(2) + np.uint32(1)
    u = t.astype(np.float32) * U_SCALE            # exact: t < 2^24
    r = np.sqrt(F32(-2.0) * log_f32(u))
    c, s = sincos_quadrant(xb)
    return r * c, r * s


def normal4(seed, stream, row, group):
    """Four float32 normals for each (row, group); returns array of shape row.shape + (4,)."""
    row = np.asarray(row, dtype=np.uint64)
    group = np.asarray(group, dtype=np.uint32)
    row, group = np.broadcast_arrays(row, group)
    seed = int(seed)
    ctr = ((row & MASK32).astype(np.uint32), (row >> np.uint64(32)).astype(np.uint32),
           group.astype(np.uint32), np.full(row.shape, stream, dtype=np.uint32))
    key = (np.full(row.shape, seed & 0xFFFFFFFF, dtype=np.uint32),
           np.full(row.shape, (seed >> 32) & 0xFFFFFFFF, dtype=np.uint32))
    x0, x1, x2, x3 = philox4x32_10(ctr, key)
    z0, z1 = box_muller(x0, x1)
    z2, z3 = box_muller(x2, x3)
    return np.stack([z0, z1, z2, z3], axis=-1)


def round_bf16(x):
    """Round binary32 to the nearest bfloat16 (ties to even), returned as binary32."""
    b = _bits(x).astype(np.uint64)
    b = (b + np.uint64(0x7FFF) + ((b >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return b.astype(np.uint32).view(np.float32)


def omega(seed, stream, n, ell, row0=0):
    """Sketch matrix rows row0..row0+n-1, columns 0..ell-1 (float32, bf16-exact values)."""
    groups = (ell + 3) // 4
    rows = np.arange(row0, row0 + n, dtype=np.uint64)[:, None]
    g = np.arange(groups, dtype=np.uint32)[None, :]
    z = normal4(seed, stream, rows, g).reshape(n, groups * 4)[:, :ell]
    z = round_bf16(z)
    z = np.where(np.abs(z) < BF16_FLUSH, np.float32(0.0), z).astype(np.float32)
    return np.ascontiguousarray(z)


def normal_flat(seed, stream, count, start=0):
    """count float32 normals; element e comes from normal4(seed, stream, e>>2, 0)[e&3]."""
    r0 = start // 4
    r1 = (start + count + 3) // 4
    rows = np.arange(r0, r1, dtype=np.uint64)
    z = normal4(seed, stream, rows, np.zeros(rows.shape, dtype=np.uint32)).reshape(-1)
    off = start - 4 * r0
    return np.ascontiguousarray(z[off:off + count]).astype(np.float32)
