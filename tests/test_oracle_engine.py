"""Pins of the oracle's random-number layer (restating src/langmuirCore/rand.cpp:8-124).

* MT19937: the C++11 known answer (10000th output of the default seed 5489 is
  4123659995, [rand.predef]), numpy's independent legacy MT19937 (RandomState
  with an integer seed uses the same init_genrand recurrence), and libstdc++'s
  std::mt19937 compiled on the spot.
* uniform_int (Boost >= 1.47 bucket rule, the reference's third-party
  dependency, absent from the tree): properties that fix the rule -- power-of-two
  ranges take the top bits, the rejection boundary is (range+1)*bucket, an empty
  range (lo == hi) consumes no engine output, and uniformity.
* uniform_01 is u32 * 2^-32 by definition of the restatement; metropolisWithCoupling
  draws exactly once per call (rand.cpp:93).
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest


def test_mt19937_known_answer(oracle):
    assert oracle.mt_outputs(5489, 1, skip=9999)[0] == 4123659995


@pytest.mark.parametrize("seed", [1, 42, 5489, 123456789, 0xFFFFFFFF])
def test_mt19937_matches_numpy_legacy(oracle, seed):
    rs = np.random.RandomState(seed)
    key = rs.get_state()[1].astype(np.uint32)
    # the legacy generator's key after seeding is the init_genrand state x[0..623]
    from langmuir_amd.host import mt19937_seed_window
    assert np.array_equal(key, mt19937_seed_window(seed))
    # and its 32-bit draws are the tempered outputs
    out = oracle.mt_outputs(seed, 2000)
    draws = rs.randint(0, 2 ** 32, size=2000, dtype=np.uint64)
    assert np.array_equal(out.astype(np.uint64), draws)


def test_mt19937_matches_libstdcxx(oracle):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = r"""
#include <random>
#include <cstdio>
int main(){ for (unsigned s : {1u, 7u, 5489u, 2024u}) { std::mt19937 g(s); g.discard(777);
  for (int i = 0; i < 5; ++i) std::printf("%u\n", (unsigned)g()); } }
"""
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "m.cpp"), "w") as f:
            f.write(src)
        subprocess.run([gxx, "-O1", "-o", os.path.join(d, "m"), os.path.join(d, "m.cpp")], check=True)
        vals = [int(x) for x in subprocess.run([os.path.join(d, "m")], capture_output=True, text=True,
                                                check=True).stdout.split()]
    got = []
    for s in (1, 7, 5489, 2024):
        got += [int(v) for v in oracle.mt_outputs(s, 5, skip=777)]
    assert got == vals


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8])
def test_uniform_int_power_of_two_takes_top_bits(oracle, k):
    n = 1 << k
    words = oracle.mt_outputs(99, 500)
    got, used = oracle.uniform_int_words(words, 0, n - 1, 500)
    assert used == 500
    assert np.array_equal(got, (words >> (32 - k)).astype(np.int32))


@pytest.mark.parametrize("m", [3, 5, 6, 7, 10, 26, 33, 1000, 65536 * 3 + 1])
def test_uniform_int_bucket_rule(oracle, m):
    """result = u // bucket with bucket = floor((2^32-1)/m) (+1 when (2^32-1) % m == m-1);
    draws u >= m*bucket are rejected and redrawn."""
    bucket = 0xFFFFFFFF // m + (1 if 0xFFFFFFFF % m == m - 1 else 0)
    limit = m * bucket
    ok = [0, 1, bucket - 1, bucket, limit - 1]
    words = np.array(ok + [0xFFFFFFFF] * (1 if limit <= 0xFFFFFFFF else 0) + [7], dtype=np.uint32)
    got, used = oracle.uniform_int_words(words, 10, 10 + m - 1, len(ok))
    assert list(got) == [10 + u // bucket for u in ok]
    assert used == len(ok)
    if limit <= 0xFFFFFFFF:   # the rejected word is skipped and the next one used
        got2, used2 = oracle.uniform_int_words(np.array([0xFFFFFFFF, 7], np.uint32), 0, m - 1, 1)
        assert used2 == 2 and got2[0] == 7 // bucket


def test_uniform_int_empty_range_draws_nothing(oracle):
    words = np.array([5, 6, 7], dtype=np.uint32)
    got, used = oracle.uniform_int_words(words, 4, 4, 3)
    assert list(got) == [4, 4, 4] and used == 0


def test_uniform_int_is_uniform(oracle):
    x = oracle.uniform_int(12345, 0, 5, 60000)
    counts = np.bincount(x, minlength=6)
    chi2 = ((counts - 10000.0) ** 2 / 10000.0).sum()
    assert chi2 < 25.0   # 5 dof, p ~ 1e-4
