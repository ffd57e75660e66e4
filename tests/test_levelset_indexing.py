"""CPU simulation of the batch route's level-set indexing (speedy_kernels.hip:
make_grid, ws_layout's partial region, levelset_body's writes, chunk_share /
chunk_word / collect_chunks reads, divq).

For every grid below it shows that
  * each word a forward or forward + VJP block reads from the level-set
    partials lies inside the workspace's partial region, and
  * it is a word the same call's level-set pass writes (so a workspace reused
    across calls, graph-captured or not, never feeds a stale or out-of-range
    word into S), and
  * the device's multiply-high division (divq) equals integer division for
    every argument these formulas pass it.
The formulas are restated from the kernel source: this is a bounds check of
that arithmetic, not a parity pin."""
import numpy as np
import pytest

BLOCK, CHUNK_LOADS, MAX_J = 256, 4, 16384


def divq(n, d):
    """Device divq: q = umulhi(n, floor(2^32 / d)), one correction step."""
    if d == 1:
        return n
    q = (n * ((1 << 32) // d)) >> 32
    if n - q * d >= d:
        q += 1
    return q


def make_grid(nlon, nlat, nmember):
    ncol = nlon * nlat
    ntot = ncol * nmember
    nblk = (ntot + BLOCK - 1) // BLOCK
    lsb = min(nblk, 512)
    chunk = ((ntot + lsb - 1) // lsb + BLOCK - 1) // BLOCK * BLOCK
    ls_blocks = (ntot + chunk - 1) // chunk
    jmax = (chunk - 1) // ncol + 2
    return ncol, ntot, nblk, ls_blocks, chunk, jmax


SHAPES = [(64, 32, 1), (96, 48, 1), (128, 64, 1), (256, 128, 1), (1280, 640, 1), (96, 48, 64),
          (256, 128, 3), (256, 128, 8), (8, 4, 6), (16, 8, 5), (64, 32, 7), (4, 2, 1), (3, 1, 1000),
          (1, 1, 4095), (5, 7, 333), (640, 320, 3), (100, 50, 37), (512, 256, 2), (17, 3, 129)]


@pytest.mark.parametrize("shape", SHAPES)
def test_batch_route_level_set_words_are_in_bounds_and_written_by_the_call(shape):
    nlon, nlat, nmember = shape
    ncol, ntot, nblk, ls_blocks, chunk, jmax = make_grid(nlon, nlat, nmember)
    assert ls_blocks <= CHUNK_LOADS * BLOCK and jmax <= MAX_J   # the host's capacity checks
    npart = max(2 * nblk, ls_blocks * jmax)                      # ws_layout
    # levelset_body: block bid writes words bid*jmax + i for the nm members its chunk touches
    nm_of = np.zeros(ls_blocks, np.int64)
    first_of = np.zeros(ls_blocks, np.int64)
    for bid in range(ls_blocks):
        c_begin, c_end = bid * chunk, min(bid * chunk + chunk, ntot)
        first_of[bid] = divq(c_begin, ncol)
        nm_of[bid] = divq(c_end - 1, ncol) - first_of[bid] + 1
        assert 1 <= nm_of[bid] <= jmax
        assert bid * jmax + nm_of[bid] - 1 < npart
    # readers: member m ORs the words of chunks b0(m)..b1(m) (chunk_share clamps to b1,
    # chunk_word reads b0 and b1, collect_chunks walks the range)
    for m in range(nmember):
        b0, b1 = divq(m * ncol, chunk), divq(m * ncol + ncol - 1, chunk)
        assert b0 == (m * ncol) // chunk and b1 == (m * ncol + ncol - 1) // chunk
        assert b1 - b0 + 1 <= CHUNK_LOADS * BLOCK
        for b in range(b0, b1 + 1):
            first = divq(b * chunk, ncol)
            assert first == (b * chunk) // ncol == first_of[b]
            j = m - first
            assert 0 <= j < nm_of[b], (m, b)          # written by this call's pass
            assert b * jmax + j < npart
    # forward blocks: the block's first member and the member after it (nm > 1)
    for bid in range(nblk):
        blk0 = bid * BLOCK
        m_first = divq(blk0, ncol)
        nm = divq(min(blk0 + BLOCK, ntot) - 1, ncol) - m_first + 1
        assert m_first + nm - 1 < nmember
        if ncol >= BLOCK:
            assert nm <= 2


def test_divq_matches_integer_division():
    rng = np.random.default_rng(0)
    for d in [1, 2, 3, 7, 255, 256, 257, 3072, 32768, 65536, 819200, 1 << 20, (1 << 31) - 1]:
        ns = list(rng.integers(0, 1 << 32, 2000, dtype=np.uint64)) + [0, d - 1, d, d + 1, (1 << 32) - 1]
        for n in ns:
            n = int(n)
            assert divq(n, d) == n // d, (n, d)
