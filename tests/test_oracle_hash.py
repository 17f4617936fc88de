"""Pins the oracle's C++ hashes (oracle_hash64_2, oracle_hash64) without reading their code:

  * tests/golden/hash_vectors.json holds known answers evaluated from the constants, shift
    amounts and statement order parsed out of the reference's own header
    (cuda/utiles_gpu.cuh:12-31, by tests/golden/make_fixtures.py);
  * hash64_2 is SplitMix64's output function, so hash64_2(seed + i * 0x9E3779B97F4A7C15)
    for i = 1, 2, ... must reproduce the published SplitMix64 output stream (Steele, Lea,
    Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014; reference C code
    splitmix64.c by S. Vigna): seed 1234567 gives 6457827717110365317,
    3203168211198807973, 9817491932198370423, 4593380528125082431, 16408922859458223821,
    and seed 0 gives 0xE220A8397B1DCDAF first.
A dropped xor-shift, a swapped multiplier or a wrong shift amount fails both checks.
"""
import json
import os

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1
GOLDEN_GAMMA = 0x9E3779B97F4A7C15


def _vectors():
    return json.load(open(os.path.join(HERE, "golden", "hash_vectors.json")))


def test_cpp_hash64_2_matches_reference_vectors():
    d = _vectors()
    for x, y in zip(d["inputs"], d["hash64_2"]):
        assert O.hash64_2(int(x)) == int(y), x


def test_cpp_hash64_matches_reference_vectors():
    d = _vectors()
    for x, y in zip(d["inputs"], d["hash64"]):
        assert O.hash64(int(x)) == int(y), x


def test_hash64_2_is_the_published_splitmix64_stream():
    published = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                 4593380528125082431, 16408922859458223821]
    state = 1234567
    for want in published:
        state = (state + GOLDEN_GAMMA) & M64
        assert O.hash64_2(state) == want
    assert O.hash64_2(GOLDEN_GAMMA) == 0xE220A8397B1DCDAF


def test_hash64_is_a_bijection_on_a_sample():
    # every step of hash64 is invertible (odd multipliers, xor-shifts), so no collisions
    xs = list(range(20000)) + [M64 - i for i in range(20000)]
    assert len({O.hash64(x) for x in xs}) == len(xs)
