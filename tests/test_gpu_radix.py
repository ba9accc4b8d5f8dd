"""GPU parity of the partitioned joins (RHO, RHT, RJ, RHO_atomic) against the
oracle: exact counts, CHT checksum and sorted (R.payload, S.payload) pairs."""
import numpy as np
import pytest

import gpu_util as gu

pytestmark = pytest.mark.gpu

RADIX_ALGOS = ["RHO", "RHT", "RJ", "RHO_atomic"]


def rel(keys, payload=None):
    keys = np.asarray(keys, np.uint32)
    out = np.empty((len(keys), 2), np.uint32)
    out[:, 0] = keys
    out[:, 1] = np.arange(len(keys), dtype=np.uint32) if payload is None else payload
    return out


def assert_join(tbj, oracle, algo, R, S, **opts):
    res, got = gu.run(tbj, algo, R, S, cap=None, **opts)
    n, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    assert res.matches == n
    assert res.pairs_written == min(n, S.shape[0])
    if n <= S.shape[0]:
        ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
        assert np.array_equal(ga, ra) and np.array_equal(gb, sa)
    assert res.checksum == gu.checksum(ra, sa)
    return res


@pytest.mark.parametrize("algo", RADIX_ALGOS)
def test_reference_generators_small(tbj, oracle, algo):
    nR, nS = 1 << 16, 1 << 18
    R = tbj.gen_pk(11111, nR)
    for S in (tbj.gen_fk(22222, nS, nR), tbj.gen_fk_sel(22222, nS, 2 * nR),
              tbj.gen_zipf(22222, nS, nR, 1.0)):
        assert_join(tbj, oracle, algo, R, S)


@pytest.mark.parametrize("passes", [1, 2])
@pytest.mark.parametrize("bits", [4, 9, 13, 16])
def test_pass_shapes(tbj, oracle, passes, bits):
    nR, nS = 100003, 300007
    R = tbj.gen_pk(7, nR)
    S = tbj.gen_zipf(8, nS, nR, 0.8)
    if passes == 1 and bits > 11:
        pytest.skip("one pass holds at most 11 bits")
    res = assert_join(tbj, oracle, "RHO", R, S, radix_bits=bits, passes=passes)
    assert res.passes == passes and res.radix_bits == bits


# every join family that partitions: bucket chains, histogram ranges, and the
# LDS sort with its bucket-directory search (RSM radix, PSM learned ranges,
# MWAY = RSM's kernels with the bounded merge)
PART_ALGOS = ["RHO", "RHT", "RSM", "PSM", "MWAY"]


@pytest.mark.parametrize("algo", PART_ALGOS)
def test_duplicates_and_extreme_keys(tbj, oracle, algo):
    # keys 0 and 0xFFFFFFFF make a sort task's key span 2^32 (the edge of the
    # sort bucket's exact floor)
    rng = np.random.default_rng(1)
    rk = np.concatenate([rng.integers(0, 5000, 20000), [0, 0, 0xFFFFFFFF, 0xFFFFFFFE]])
    sk = np.concatenate([rng.integers(0, 6000, 30000), [0, 0xFFFFFFFF, 0xFFFFFFFF]])
    R, S = rel(rk), rel(sk)
    n = oracle.count(R[:, 0], S[:, 0])
    res, got = gu.run(tbj, algo, R, S, cap=n)
    _, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
    assert res.matches == n and np.array_equal(ga, ra) and np.array_equal(gb, sa)


@pytest.mark.parametrize("algo", PART_ALGOS)
def test_one_partition_spanning_key_range(tbj, oracle, algo):
    # a single task whose R range holds 0 and 0xFFFFFFFF (span 2^32) and a
    # repeated key (a sort bucket far above 32 entries: the bitonic path)
    rk = np.concatenate([[0, 0xFFFFFFFF, 7, 7], np.full(300, 12345), np.arange(100, 1100)])
    sk = np.concatenate([[0, 0xFFFFFFFF, 0xFFFFFFFF, 7], np.full(5, 12345),
                         np.arange(50, 1500)])
    assert_join(tbj, oracle, algo, rel(rk), rel(sk))


@pytest.mark.parametrize("algo", PART_ALGOS)
@pytest.mark.parametrize("ns", [1, 4095, 4096, 4097, 16384, 16385, 40000])
def test_materialising_join_sub_chunk_edges(tbj, oracle, algo, ns):
    # one partition (|R| below one table): the task list splits S into tasks of
    # up to four sub-chunks of 4096 tuples, whose loads run ahead of the claims
    rng = np.random.default_rng(ns)
    R = rel(rng.integers(0, 1500, 2000))  # repeated R keys: several matches per S tuple
    S = rel(rng.integers(0, 1800, ns))
    n, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    res, got = gu.run(tbj, algo, R, S, cap=n)
    assert res.matches == n and res.pairs_written == n
    ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
    assert np.array_equal(ga, ra) and np.array_equal(gb, sa)
    assert res.checksum == gu.checksum(ra, sa)


@pytest.mark.parametrize("algo", PART_ALGOS)
def test_oversized_partition_is_chunked(tbj, oracle, algo):
    # one key repeated far beyond the LDS table capacity on the build side:
    # its partition splits into several R builds (and, for the sort join, a
    # bucket of 4096 equal keys goes through the bitonic path)
    rk = np.concatenate([np.full(10000, 42), np.arange(1, 3000)])
    sk = np.concatenate([np.full(50, 42), np.arange(1, 9000)])
    R, S = rel(rk), rel(sk)
    n = oracle.count(R[:, 0], S[:, 0])
    res, got = gu.run(tbj, algo, R, S, cap=n)
    _, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
    assert res.matches == n and res.max_r_partition > 4096
    assert np.array_equal(ga, ra) and np.array_equal(gb, sa)


@pytest.mark.parametrize("algo", ["RHO", "RHT", "RSM"])
def test_tasks_beyond_the_task_list(tbj, oracle, algo):
    # The host sizes the task list without the count.  One partition whose R
    # side spans 12 builds and whose S side 12 chunks holds 144 tasks, more
    # than the list (partitions + R builds + S chunks + 64 = 32 + 12 + 12 + 64):
    # the join decodes the rest from the per-partition task offsets.
    nr, ns = 12 * 4096, 12 * 1024
    R, S = rel(np.full(nr, 42)), rel(np.full(ns, 42))
    res, _ = gu.run(tbj, algo, R, S, materialize=False, s_chunk=1024, radix_bits=5)
    assert res.matches == nr * ns
    assert res.checksum == (ns * (nr * (nr - 1) // 2) + nr * (ns * (ns - 1) // 2)) % (1 << 64)


def test_empty_and_tiny(tbj, oracle):
    R, S = rel([5]), rel([5, 5, 6])
    res, got = gu.run(tbj, "RHO", R, S)
    assert res.matches == 2 and sorted(got[:, 1].tolist()) == [0, 1]
    res, _ = gu.run(tbj, "RHO", rel([]), S, materialize=False)
    assert res.matches == 0
    res, _ = gu.run(tbj, "RHO", R, rel([]), materialize=False)
    assert res.matches == 0


def test_output_capacity_truncates_but_counts(tbj, oracle):
    nR, nS = 5000, 20000
    R = tbj.gen_pk(3, nR)
    S = tbj.gen_fk(4, nS, nR)
    res, got = gu.run(tbj, "RHO", R, S, cap=1234)
    assert res.matches == nS and res.pairs_written == 1234
    assert np.all(R[got[:, 0], 0] == S[got[:, 1], 0])


@pytest.mark.parametrize("z", [0.0, 1.0, 1.5])
def test_learned_partitioner_matches_radix(tbj, oracle, z):
    nR, nS = 1 << 18, 1 << 20
    R = tbj.gen_pk(11111, nR)
    S = tbj.gen_zipf(22222, nS, nR, z) if z else tbj.gen_fk(22222, nS, nR)
    S = tbj.remap_odd(S, nR)  # half of S outside R's key range
    a = assert_join(tbj, oracle, "RHO", R, S, partitioner="learned")
    b = assert_join(tbj, oracle, "RHO", R, S, partitioner="radix")
    assert a.partitioner_used == tbj.PART_LEARNED and a.fallback == 0
    # R's keys 1..|R| are uniform: the fit is the linear map of the key range
    assert a.learned_model == 2 and b.learned_model == 0
    assert a.s_dropped == nS // 2 and b.s_dropped == 0
    assert (a.matches, a.checksum) == (b.matches, b.checksum)


def skewed_unique_keys(rng, n):
    """Unique uint32 keys whose CDF is far from a line: half of them dense at
    the bottom of the range, half spread thinly above (gaps of 500..1500)."""
    dense = np.arange(1, n // 2 + 1, dtype=np.uint64)
    sparse = np.uint64(n) + np.cumsum(rng.integers(500, 1500, n - n // 2)).astype(np.uint64)
    k = np.concatenate([dense, sparse])
    assert int(k.max()) < 2**32
    return rng.permutation(k).astype(np.uint32)


@pytest.mark.parametrize("sampled", [True, -1])
def test_learned_piecewise_model_on_skewed_keys(tbj, oracle, sampled):
    # a sample that is not near-linear: the knots-and-cells model, in the
    # sample-sized flow (R's scatter measures the range S drops against) and
    # in the exact flow (the exact range read beside the sample's sort)
    rng = np.random.default_rng(21)
    nR = 1 << 18
    rk = skewed_unique_keys(rng, nR)
    sk = np.concatenate([rng.choice(rk, 1 << 19),
                         rng.integers(0, 2**32, 1 << 17, dtype=np.uint64).astype(np.uint32)])
    res = assert_join(tbj, oracle, "RHO", rel(rk), rel(sk[rng.permutation(len(sk))]),
                      partitioner="learned", sampled_regions=sampled, radix_bits=14, passes=2)
    assert res.partitioner_used == tbj.PART_LEARNED and res.learned_model == 1
    assert res.redone_exact == 0


@pytest.mark.parametrize("where", ["bottom", "top"])
@pytest.mark.parametrize("sampled", [True, -1])
def test_learned_linear_model_at_range_ends(tbj, oracle, where, sampled):
    # the linear model's key range widened past the sample's extremes, cut
    # at 0 and at 2^32 - 1; S keys at the very ends and just outside R's keys
    rng = np.random.default_rng(22)
    nR = 200000
    base = 0 if where == "bottom" else 2**32 - 2**21
    rk = (base + rng.choice(2**21, nR, replace=False)).astype(np.uint32)
    ends = np.array([0, 1, 2**32 - 1, 2**32 - 2, base, rk.min(), rk.max()], np.uint64)
    sk = np.concatenate([rng.choice(rk, 600000).astype(np.uint64), ends,
                         ends.astype(np.int64).clip(1) - 1]).astype(np.uint32)
    res = assert_join(tbj, oracle, "RHO", rel(rk), rel(sk[rng.permutation(len(sk))]),
                      partitioner="learned", sampled_regions=sampled, radix_bits=12, passes=2)
    assert res.partitioner_used == tbj.PART_LEARNED and res.learned_model == 2


def test_learned_fallback_to_radix(tbj, oracle):
    nR, nS = 1 << 16, 1 << 18
    R = tbj.gen_pk(5, nR)
    S = tbj.gen_fk(6, nS, nR)
    res = assert_join(tbj, oracle, "RHO", R, S, partitioner="learned", learned_bound=16)
    assert res.fallback == 1 and res.partitioner_used == tbj.PART_RADIX


def test_reference_config1_count(tbj):
    # |R| = |S| = 2^21, default generators: the reference prints Matches = 2097152
    R = tbj.gen_pk(11111, 1 << 21, payload=tbj.PAYLOAD_ZERO)
    S = tbj.gen_fk(22222, 1 << 21, 1 << 21, payload=tbj.PAYLOAD_ZERO)
    res, _ = gu.run(tbj, "RHO", R, S, materialize=False)
    assert res.matches == 2097152 and res.checksum == 0


def sampled_positions(n):
    """Tuples the sample-sized passes sample: exactly one 16-tuple group in
    every block of 256 tuples, at a hashed position."""
    i = np.arange(n, dtype=np.uint64)
    block = i >> np.uint64(8)
    pick = ((block * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)) >> np.uint64(28)
    return ((i >> np.uint64(4)) & np.uint64(15)) == pick


@pytest.mark.parametrize("algo", ["RHO", "RHT", "RSM", "PSM"])
@pytest.mark.parametrize("z", [0.0, 1.2])
def test_sampled_passes(tbj, oracle, algo, z):
    nR, nS = 1 << 18, 1 << 20
    R = tbj.gen_pk(11111, nR)
    S = tbj.gen_zipf(22222, nS, nR, z) if z else tbj.gen_fk(22222, nS, nR)
    res = assert_join(tbj, oracle, algo, R, S, sampled_regions=True, radix_bits=14, passes=2)
    assert res.passes == 2 and res.redone_exact == 0


@pytest.mark.parametrize("nR,nS", [(4097, 1), (100003, 300007), (1 << 20, 123457), (1 << 16, 1 << 22)])
def test_sampled_passes_ragged_rows(tbj, oracle, nR, nS):
    # relations whose tiles do not fill the XCD eighths or the rows evenly
    R = tbj.gen_pk(5, nR)
    S = tbj.gen_zipf(6, nS, nR, 0.7)
    res = assert_join(tbj, oracle, "RHO", R, S, sampled_regions=True, passes=2,
                      radix_bits=max(9, min(16, int(np.ceil(np.log2(nR / 512))))))
    assert res.passes == 2 and res.redone_exact == 0


def test_sampled_learned_drops_outside_range(tbj, oracle):
    nR, nS = 1 << 18, 1 << 20
    R = tbj.gen_pk(11111, nR)
    S = tbj.remap_odd(tbj.gen_zipf(22222, nS, nR, 1.0), nR)
    res = assert_join(tbj, oracle, "RHO", R, S, partitioner="learned", sampled_regions=True,
                      radix_bits=14, passes=2)
    assert res.partitioner_used == tbj.PART_LEARNED and res.s_dropped == nS // 2
    assert res.redone_exact == 0


@pytest.mark.parametrize("algo", ["RHO", "RHT", "RSM"])
def test_dense_keys_narrow_the_second_pass(tbj, oracle, algo):
    # R's keys 1..|R| span exactly |R| values: the plan one bit narrower (9
    # instead of 10 bits here) holds exactly |R| / 2^9 = 4096 R tuples, one
    # build, per partition
    nR, nS = 1 << 21, 1 << 23
    R = tbj.gen_pk(11111, nR)
    S = tbj.gen_fk(22222, nS, nR)
    res = assert_join(tbj, oracle, algo, R, S, sampled_regions=True)
    assert (res.passes, res.radix_bits, res.num_partitions) == (2, 9, 1 << 9)
    assert res.max_r_partition == 4096 and res.redone_exact == 0
    # key 1 moved to |R| + 2 keeps the plan (the range [2, |R| + 2] spans
    # |R| + 1 values)
    R2 = R.copy()
    R2[R2[:, 0] == 1, 0] = nR + 2
    res = assert_join(tbj, oracle, algo, R2, S, sampled_regions=True)
    assert (res.passes, res.radix_bits) == (2, 10)
    # asked-for bits are kept
    res = assert_join(tbj, oracle, algo, R, S, sampled_regions=True, radix_bits=10)
    assert res.radix_bits == 10


def test_sample_sized_region_overflow_redoes_exact(tbj, oracle):
    # tuples the sample sees all carry key 1; the rest carry key 2, whose
    # regions the sample believes empty
    nR, nS = 1 << 14, 1 << 16
    R = tbj.gen_pk(1, nR)
    S = rel(np.where(sampled_positions(nS), 1, 2))
    res = assert_join(tbj, oracle, "RHO", R, S, radix_bits=12, passes=2, sampled_regions=True)
    assert res.redone_exact == 1 and res.matches == nS


def test_exact_passes_when_asked(tbj, oracle):
    nR, nS = 1 << 18, 1 << 20
    R = tbj.gen_pk(3, nR)
    S = tbj.gen_fk(4, nS, nR)
    res = assert_join(tbj, oracle, "RHO", R, S, sampled_regions=-1, radix_bits=14, passes=2)
    assert res.redone_exact == 0


def test_profile_phases(tbj, oracle):
    nR, nS = 1 << 18, 1 << 20
    R = tbj.gen_pk(3, nR)
    S = tbj.gen_fk(4, nS, nR)
    res = assert_join(tbj, oracle, "RHO", R, S, profile=True, radix_bits=14, passes=2)
    ph = res.as_dict()["phase_ms"]
    assert ph["join"] > 0 and ph["scatter1"] > 0 and ph["scatter2"] > 0


@pytest.mark.parametrize("part_tiles", [1, 31, 32, 33, 65])
def test_segments_across_tile_groups(tbj, oracle, part_tiles):
    # Output offsets come from sums over groups of 32 consecutive tiles; a
    # second-pass segment (one first-pass partition) of about part_tiles
    # tiles of 4096 tuples then starts and ends at varying places inside
    # groups, and groups hold pieces of several segments.  16 first-pass
    # partitions (4 + 4 bits); a skewed key mix leaves some partitions
    # almost empty and others several groups long.
    rng = np.random.default_rng(100 + part_tiles)
    n = 16 * part_tiles * 4096
    nR = n // 4
    rk = rng.integers(0, 1 << 20, nR, dtype=np.uint32)
    # bin p (low 4 bits) gets a share proportional to p + 1 of S's keys
    w = np.arange(1, 17, dtype=np.float64)
    bins = rng.choice(16, n - nR, p=w / w.sum())
    sk = ((rng.integers(0, 1 << 16, n - nR, dtype=np.uint32) << 4) | bins).astype(np.uint32)
    sk[::3] = rk[rng.integers(0, nR, sk[::3].shape[0])]
    assert_join(tbj, oracle, "RHO", rel(rk), rel(sk), radix_bits=8, passes=2)
