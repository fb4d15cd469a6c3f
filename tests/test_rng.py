"""The package's SeededRng (bench inputs) matches the pinned reference sequence
and the oracle's restatement draw for draw."""
import numpy as np

import oracle
from gprx_amd.rng import SeededRng
from test_oracle_exact import SEED0, SEED42


def test_pinned():
    r = SeededRng(0)
    assert [r.next_u64() for _ in range(4)] == SEED0
    r = SeededRng(42)
    assert [r.next_u64() for _ in range(4)] == SEED42


def test_matches_oracle_stream():
    a, b = SeededRng(2), oracle.Rng(2)
    assert np.array_equal(a.units(500), b.unit(500))
    assert np.array_equal(a.normals(500), b.normal(500))
    for mx in [0, 1, 5, 1000, 2**40]:
        assert a.up_to(mx) == b.up_to(mx)
    i1 = np.arange(100, dtype=np.uint64)
    i2 = i1.copy()
    a.shuffle(i1)
    b.shuffle(i2)
    assert np.array_equal(i1, i2)


def test_native_bulk_draws_match_python_stream():
    """libgprx_rng.so continues the same stream as the Python generator:
    interleaved bulk and scalar draws agree bit for bit."""
    import gprx_amd.rng as grng

    if grng._native() is None:
        import pytest

        pytest.skip("libgprx_rng.so not built")
    a, b = SeededRng(3), SeededRng(3)
    for n in (1000, 77, 4096):
        assert np.array_equal(a.units(n), b.units(n, native=False))
        assert np.array_equal(a.normals(n), b.normals(n, native=False))
        assert a.next_u64() == b.next_u64()


# ---------------------------------------------------------------------------
# Closed forms from the pinned first words (src/rng.rs:98-120).  Each check
# is worked from the words themselves, not from the generator's code, and is
# chosen so that a plausible restatement mistake fails it (see
# test_closed_forms_reject_plausible_mistakes).
# ---------------------------------------------------------------------------
import math


def _generators():
    """The oracle's generator and the package's (both must hold the pins)."""
    return [oracle.Rng, SeededRng]


def _unit(gen, seed, k):
    r = gen(seed)
    return [float(r.unit()) if gen is oracle.Rng else r.unit() for _ in range(k)]


def test_unit_is_top_53_bits_times_2_pow_minus_53():
    # unit() = (w >> 11) * 2^-53 (src/rng.rs:45-48); exact in f64, so equality.
    for gen in _generators():
        for seed, words in ((0, SEED0), (42, SEED42)):
            got = _unit(gen, seed, 4)
            want = [(w >> 11) * 2.0 ** -53 for w in words]
            assert got == want, (gen, seed)
    # The worked value: 0x53175d61490b23df >> 11 = 2923514112319844, / 2^53.
    assert _unit(oracle.Rng, 0, 1)[0] == 2923514112319844 / 9007199254740992


def test_up_to_power_of_two_range_is_top_bits():
    # range 2^k divides 2^64, so Lemire's threshold (2^64 - 2^k) mod 2^k is 0,
    # nothing is rejected and (w * 2^k) >> 64 = w >> (64 - k): the top k bits.
    # Top 3 bits of the seed-0 words 0x5317.., 0x61da.., 0x5c0f.., 0x02ee..:
    # 010, 011, 010, 000.
    for gen in _generators():
        r = gen(0)
        assert [r.up_to(7) for _ in range(4)] == [2, 3, 2, 0]
        r = gen(42)  # 0xd076.. 0x519e.. 0xfbe0.. 0xb37d..: top bit 1, 0, 1, 1
        assert [r.up_to(1) for _ in range(4)] == [1, 0, 1, 1]


def test_up_to_three_values_is_floor_of_three_units():
    # range 3: threshold = (2^64 - 3) mod 3 = 1; a word is rejected only when
    # 3w mod 2^64 = 0, which none of these are.  Then up_to(2) = floor(3 w / 2^64):
    # seed 0 words are 0.3246, 0.3822 of 2^64 -> 0, 1; seed 42's 0.8143 -> 2.
    for gen in _generators():
        r = gen(0)
        assert [r.up_to(2), r.up_to(2)] == [0, 1]
        assert gen(42).up_to(2) == 2


def test_fisher_yates_worked_by_hand():
    # shuffle_indices (src/svgp/factor/adam.rs:85-90): for i = len-1 down to
    # 1, j = up_to(i), swap(i, j).
    # seed 0, [0,1,2]: i=2 j=floor(3*0.3246)=0 -> [2,1,0]; i=1 j=top bit of
    #   0x61da..=0 -> [1,2,0].
    # seed 42, [0,1,2]: i=2 j=floor(3*0.8143)=2 -> [0,1,2]; i=1 j=top bit of
    #   0x519e..=0 -> [1,0,2].
    # seed 0, [0,1,2,3]: i=3 j=top 2 bits of 0x5317..=1 -> [0,3,2,1]; i=2
    #   j=floor(3*0.3822)=1 -> [0,2,3,1]; i=1 j=top bit of 0x5c0f..=0 -> [2,0,3,1].
    # seed 0, [0,1]: j = top bit 0 -> [1,0]; seed 42, [0,1]: j = 1 -> [0,1].
    cases = {(0, 2): [1, 0], (42, 2): [0, 1], (0, 3): [1, 2, 0], (42, 3): [1, 0, 2],
             (0, 4): [2, 0, 3, 1]}
    for gen in _generators():
        for (seed, n), want in cases.items():
            idx = np.arange(n, dtype=np.uint64)
            gen(seed).shuffle(idx)
            assert idx.tolist() == want, (gen, seed, n)


def test_box_muller_from_two_pinned_words():
    # unit_normal = sqrt(-2 ln u1) cos(2 pi u2) (src/rng.rs:81-87), u1 and u2
    # the first two open units; for seed 0 these are 0.32457526803140670 and
    # 0.38223929651167343, which give -1.1079085986338315.
    u1 = (SEED0[0] >> 11) * 2.0 ** -53
    u2 = (SEED0[1] >> 11) * 2.0 ** -53
    want = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
    assert abs(want - (-1.1079085986338315)) < 1e-15
    assert abs(float(oracle.Rng(0).normal(1)[0]) - want) <= 1e-15
    assert abs(SeededRng(0).normals(1)[0] - want) <= 1e-15
    u1 = (SEED42[0] >> 11) * 2.0 ** -53
    u2 = (SEED42[1] >> 11) * 2.0 ** -53
    want = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
    assert abs(float(oracle.Rng(42).normal(1)[0]) - want) <= 1e-15


def test_closed_forms_reject_plausible_mistakes():
    """The expected values above are not met by the usual restatement
    mistakes, applied to the same pinned words."""
    w0, w42 = SEED0, SEED42
    # unit with >> 12 halves the value
    assert (w0[0] >> 12) * 2.0 ** -53 != (w0[0] >> 11) * 2.0 ** -53
    # up_to by modulo: low bits instead of top bits
    assert [w % 8 for w in w0] != [2, 3, 2, 0]
    assert [w0[0] % 3, w0[1] % 3] != [0, 1]

    # Fisher-Yates with an off-by-one bound (j = up_to(i - 1)) or run forward
    def fy(words, n, bound_shift=0, forward=False):
        it = iter(words)

        def up_to(mx):
            # floor(w (mx+1) / 2^64); none of these words is rejected
            return (next(it) * (mx + 1)) >> 64

        idx = list(range(n))
        order = range(1, n) if forward else range(n - 1, 0, -1)
        for i in order:
            j = up_to(i + bound_shift)
            idx[i], idx[j] = idx[j], idx[i]
        return idx

    assert fy(w42, 3) == [1, 0, 2] and fy(w0, 4) == [2, 0, 3, 1]  # the rule itself
    assert fy(w42, 3, bound_shift=-1) != [1, 0, 2]
    assert fy(w0, 4, forward=True) != [2, 0, 3, 1]
    # Box-Muller with sin, or with u1 and u2 exchanged
    u1 = (w0[0] >> 11) * 2.0 ** -53
    u2 = (w0[1] >> 11) * 2.0 ** -53
    good = math.sqrt(-2 * math.log(u1)) * math.cos(2 * math.pi * u2)
    assert abs(math.sqrt(-2 * math.log(u1)) * math.sin(2 * math.pi * u2) - good) > 0.1
    assert abs(math.sqrt(-2 * math.log(u2)) * math.cos(2 * math.pi * u1) - good) > 0.1
