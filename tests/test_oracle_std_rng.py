"""Pins the oracle's restatement of GsInit::Random's generator
(crates/optics-inverse/src/gs.rs:449-458: rand 0.8 StdRng::seed_from_u64,
Uniform::new(0.0, span), lo + sample).

- The ChaCha block function against RFC 7539's published ChaCha20 answers
  (the section 2.3.2 block and the all-zero key stream of appendix A.1),
  which fix the constants, the quarter round, the round order and the final
  addition; StdRng runs the same function with 6 double rounds.
- StdRng's u64 stream is the ChaCha12 blocks 0, 1, 2, ... of the
  PCG32-seeded key, two words per value, low word first (a composition
  check: the seeding itself has no published vector here, so the stream
  stays parity unpinned).
- The Uniform<f64> mapping: every phase is lo + ((u >> 12) * 2^-52) * scale
  for the stream's u, in [lo, lo + span), with span = (hi - lo) clamped to
  [0, 2 pi]; a zero span gives lo everywhere.
"""
import math

import numpy as np

import oc_types as T

def test_chacha20_rfc7539_block(orc):
    # RFC 7539 section 2.3.2: key 00..1f, counter 1, nonce 00000009 0000004a 00000000
    key = np.array([int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)], np.uint32)
    counter = 1 | (0x09000000 << 32)
    stream = 0x4A000000
    got = orc.chacha_block(key, counter, stream, 10)
    want = [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7, 0x0368C033, 0x9AAA2204, 0x4E6CD4C3,
            0x466482D2, 0x09AA9F07, 0x05D7C214, 0xA2028BD9, 0xD19C12B5, 0xB94E16DE, 0xE883D0CB, 0x4E3C50A2]
    assert [int(w) for w in got] == want


def test_chacha20_zero_key_stream(orc):
    # RFC 7539 appendix A.1, test vectors 1 and 2 (all-zero key and nonce, blocks 0 and 1)
    b0 = orc.chacha_block(np.zeros(8, np.uint32), 0, 0, 10)
    want0 = bytes.fromhex("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
                          "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    assert b0.astype("<u4").tobytes() == want0
    b1 = orc.chacha_block(np.zeros(8, np.uint32), 1, 0, 10)
    assert b1.astype("<u4").tobytes()[:16] == bytes.fromhex("9f07e7be5551387a98ba977c732d080d")


def _pcg32_key(seed):
    mask = (1 << 64) - 1
    key = []
    s = seed
    for _ in range(8):
        s = (s * 6364136223846793005 + 11634580027462260723) & mask
        xs = (((s >> 18) ^ s) >> 27) & 0xFFFFFFFF
        rot = s >> 59
        key.append(((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF)
    return np.array(key, np.uint32)


def test_std_rng_stream_is_chacha12_blocks(orc):
    for seed in (0, 3, 7, 11, 2 ** 64 - 1):
        vals = orc.std_rng_u64(seed, 40)  # five blocks
        key = _pcg32_key(seed)
        words = np.concatenate([orc.chacha_block(key, b, 0, 6) for b in range(5)]).astype(np.uint64)
        want = words[0::2] | (words[1::2] << np.uint64(32))
        assert np.array_equal(vals, want)
    assert not np.array_equal(orc.std_rng_u64(7, 8), orc.std_rng_u64(8, 8))
    assert np.array_equal(orc.std_rng_u64(7, 8), orc.std_rng_u64(7, 8))


def test_gs_random_phase_mapping(orc):
    n = 20000
    lo, hi = -math.pi, math.pi
    ph = orc.gs_random_phase(7, lo, hi, n)
    u = orc.std_rng_u64(7, n)
    span = 2 * math.pi  # (hi - lo) clamped to TAU: 2*pi rounds to TAU exactly
    v01 = (u >> np.uint64(12)).astype(np.float64) * 2.0 ** -52
    assert np.array_equal(ph, lo + (v01 * span + 0.0))
    assert ph.min() >= lo and ph.max() < lo + span
    assert abs(ph.mean() - (lo + span / 2)) < 0.05
    assert abs(ph.var() - span ** 2 / 12) < 0.05 * span ** 2 / 12
    # span clamped to 2 pi, and a zero span gives lo
    wide = orc.gs_random_phase(3, 0.0, 10.0, 1000)
    assert wide.max() < 2 * math.pi and wide.min() >= 0.0
    assert np.array_equal(orc.gs_random_phase(3, 0.5, 0.5, 16), np.full(16, 0.5))
    narrow = orc.gs_random_phase(3, 1.0, 1.5, 1000)
    assert narrow.min() >= 1.0 and narrow.max() < 1.5


def test_gs_random_init_runs_in_oracle(orc):
    """The reference's random-start FPR case (gs.rs:1005-1049): the patience
    stop fires within the iteration cap."""
    n, d, lam = 64, 10e-6, 500e-9
    x = (np.arange(n) - (n - 1) / 2) * d
    r2 = x[None, :] ** 2 + x[:, None] ** 2
    u = np.exp(-r2 / (120e-6) ** 2).astype(complex)
    target = np.where(r2 <= (180e-6) ** 2, 1.0, 0.0)
    st, sol = orc.gs_run(u, d, d, lam, 6e-3, target, target ** 2, T.AsmConfig(2.0, 0, 0, 8.0, 1.0),
                         max_iterations=120, patience=5, merit_tolerance=1e-3, variant=2, beta=0.8, init_seed=3)
    assert st == 0
    assert sol["converged"] and sol["iterations"] <= 120
