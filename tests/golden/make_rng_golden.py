"""Writes rng_vectors.json from an independent pure-Python statement of the
PCG hash (Jarzynski & Olano 2020, as spectral-core/src/rng.rs:7-40 uses it).

The first four draws of PathRng::new(0xABCD, 0x1234) are the reference's own
frozen values (rng.rs:82); the script asserts them before writing, so the
other vectors come from a restatement that reproduces the reference's known
answer.  Run: python tests/golden/make_rng_golden.py
"""
import json
from pathlib import Path

M = 0xFFFFFFFF


def pcg32(x: int) -> int:
    state = (x * 747796405 + 2891336453) & M
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M
    return ((word >> 22) ^ word) & M


def draws(key: int, seed: int, n: int):
    st = pcg32(key ^ pcg32(seed))
    out = []
    for _ in range(n):
        st = (st + 1) & M
        out.append(pcg32(st))
    return out


def main() -> None:
    frozen = [0xA1DD5847, 0x13248A2D, 0x37278DEF, 0xDE1D5B62]
    d16 = draws(0xABCD, 0x1234, 16)
    assert d16[:4] == frozen, [hex(x) for x in d16[:4]]
    vec = {
        "frozen_first4": frozen,
        "draws_abcd_1234_16": d16,
        "draws_more": [{"key": k, "seed": s, "draws": draws(k, s, 8)}
                       for k, s in [(0, 0), (7, 12345), (8, 12345), (1, 2), (0xFFFFFFFF, 42),
                                    (123456789, 0xDED)]],
        "mix2": [{"a": a, "b": b, "out": pcg32(a ^ pcg32(b))}
                 for a, b in [(0, 0), (1, 0), (0, 1), (4095, 63), (262143, 1000), (M, M)]],
        "f32_bits": [{"key": 7, "seed": 12345,
                      "f32": [((u >> 8) / 16777216.0) for u in draws(7, 12345, 8)]}],
    }
    out = Path(__file__).with_name("rng_vectors.json")
    out.write_text(json.dumps(vec, indent=1) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
