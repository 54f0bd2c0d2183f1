"""CPU checks of the batched inputs: the reference's xorshift64* stream, the Gaussian random
field sampler, the blocked vector-Laplacian values and the stacking of coefficient fields."""
import numpy as np
import pytest

from feat3_amd import problems as pb


def test_xorshift_c_matches_python_restatement():
    # two independent restatements of util/random.hpp:62-110,259-280 (C in the setup library,
    # numpy here) must give the same doubles
    for seed in (0, 1, 7, 28054777172512, 2 ** 63 + 5):
        py = pb.XorShift64Star(seed).uniform(500, -1.0, 1.0)
        c = pb.random_rhs(500, 1, 1, np.array([], np.uint32), seed).reshape(-1)
        assert np.array_equal(py, c)
        assert np.all((py >= -1.0) & (py <= 1.0))


def test_xorshift_known_first_words():
    # xorshift64* (Vigna 2016, shifts 12/25/27) from state 1: the first state is
    # 1 ^ (1 << 25) ^ ((1 ^ (1 << 25)) >> 27) = 0x2000001 (the >>27 term vanishes), and the
    # output is that state times the multiplier modulo 2^64
    r = pb.XorShift64Star(1)
    w = int(r.next_block(1)[0])
    assert w == (0x2000001 * 2685821657736338717) % (1 << 64)
    # the zero seed is replaced by the default seed
    assert np.array_equal(pb.XorShift64Star(0).next_block(4), pb.XorShift64Star(28054777172512).next_block(4))


def test_xorshift_degenerate_range_returns_lower_bound():
    assert np.all(pb.XorShift64Star(3).uniform(10, 2.0, 2.0) == 2.0)


def test_grf_covariance():
    # squared-exponential covariance sigma^2 exp(-r^2 / (2 l^2)): variance 1 and the lag-1/2/4 cell
    # correlations at h = 1/32, estimated over 300 seeded fields
    lvl, nf = 5, 300
    m = 1 << lvl
    G = np.stack([pb.gaussian_random_field(2, lvl, seed=s).reshape(m, m) for s in range(nf)])
    var = np.mean(G * G)
    assert abs(var - 1.0) < 0.06
    for lag in (1, 2, 4):
        c = np.mean(G[:, :, :-lag] * G[:, :, lag:])
        ref = np.exp(-(lag / m) ** 2 / (2 * 0.1 ** 2))
        assert abs(c - ref) < 0.06, (lag, c, ref)
    # x fastest: correlations along both axes agree (isotropy); a transposed layout keeps this,
    # so also check the mean is near zero
    cy = np.mean(G[:, :-1, :] * G[:, 1:, :])
    assert abs(cy - np.exp(-(1 / m) ** 2 / 0.02)) < 0.06
    assert abs(np.mean(G)) < 0.1


def test_block_identity_values():
    H = pb.hierarchy(2, 1, 2, 1)
    A = H.levels[0].A
    v = pb.block_identity_values(A, 3)
    assert v.shape == (A.nnz, 3, 3)
    for h in range(3):
        assert np.array_equal(v[:, h, h], A.val)
    off = v.copy()
    for h in range(3):
        off[:, h, h] = 0
    assert np.all(off == 0)


def test_stacked_values_requires_shared_graph():
    a = pb.hierarchy(2, 1, 3, 1).levels[0].A
    b = pb.hierarchy(2, 1, 4, 1).levels[0].A
    with pytest.raises(ValueError):
        pb.stacked_values([a, b])
    k = np.exp(np.stack([pb.gaussian_random_field(2, 3, seed=s) for s in range(3)]))
    Hs = pb.field_hierarchies(2, 3, k)
    s = pb.stacked_values([H.levels[0].A for H in Hs])
    assert s.shape == (Hs[0].levels[0].A.nnz, 3)
    for j in range(3):
        assert np.array_equal(s[:, j], Hs[j].levels[0].A.val)


@pytest.mark.parametrize("dim,lvl,tile", [(3, 4, 64), (2, 6, 100), (3, 3, 1), (2, 2, 1000)])
def test_tile_order_is_compact_permutation(dim, lvl, tile):
    from feat3_amd import capi
    A = pb.hierarchy(dim, 1, lvl, lvl).levels[0].A
    order = np.empty(A.rows, np.uint32)
    capi.call("feat3_setup_tile_order", A.rows, pb._p(A.row_ptr), pb._p(A.col_ind), tile, pb._p(order))
    assert np.array_equal(np.sort(order), np.arange(A.rows, dtype=np.uint32))
    if tile <= 1:
        assert np.array_equal(order, np.arange(A.rows, dtype=np.uint32))
        return
    # tiles are grown from a seed by breadth-first search: the first tile is the graph ball of
    # node 0, a cube of side about tile^(1/dim) on the Q1 grid, where `tile` consecutive rows
    # would span whole grid lines
    m = (1 << lvl) + 1
    first = np.sort(order[:tile])
    coords = np.stack([(first // m ** d) % m for d in range(dim)], axis=1)
    ext = coords.max(axis=0) - coords.min(axis=0) + 1
    assert np.prod(ext) <= 2 * tile
    assert np.all(ext <= int(np.ceil(tile ** (1.0 / dim))) + 1)


@pytest.mark.parametrize("dim,deg", [(2, 1), (2, 2), (3, 1)])
def test_device_assembly_sizes_match_the_host_setup(dim, deg):
    """The closed-form sizes of the device assembly (host-side, no GPU) equal the host setup's
    counted sizes for every level."""
    import ctypes
    from feat3_amd import capi
    for lvl in range(0, 7 if dim == 2 else 5):
        a = [ctypes.c_uint64() for _ in range(3)]
        b = [ctypes.c_uint64() for _ in range(3)]
        capi.call("feat3_setup_laplace_sizes", dim, deg, lvl, *[ctypes.byref(v) for v in a])
        capi.call("feat3_device_laplace_sizes", dim, deg, lvl, *[ctypes.byref(v) for v in b])
        assert [v.value for v in a] == [v.value for v in b], (dim, deg, lvl)
        if lvl >= 1:
            capi.call("feat3_setup_prol_sizes", dim, deg, lvl, *[ctypes.byref(v) for v in a])
            capi.call("feat3_device_prol_sizes", dim, deg, lvl, *[ctypes.byref(v) for v in b])
            assert [v.value for v in a] == [v.value for v in b], (dim, deg, lvl)
