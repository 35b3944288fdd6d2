"""Writes tests/golden/philox4x32_8_kat.json: Philox4x32 with 8 rounds evaluated by
the reference's own vendored Random123 header (src/third_party/Random123/philox.h,
the generator of src/steps/rng/r123.hpp:76), compiled from the reference sources
by `make -C oracle ref` into oracle/_ref/philox_kat.

    make -C oracle ref && python tests/golden/make_philox_kat.py
"""
import json
import random
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import bind  # noqa: E402


def cases():
    rnd = random.Random(20261203)
    out = [(0, 0, 0, 0, 0, 0), (0xFFFFFFFF,) * 6,
           (0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344, 0xA4093822, 0x299F31D0)]
    out += [tuple(rnd.getrandbits(32) for _ in range(6)) for _ in range(61)]
    # counters shaped like the engine's: (tet, period, stream, channel<<24 | block)
    out += [(t, p, r, (ch << 24) | b, 12345, 0) for t, p, r, ch, b in
            [(0, 0, 0, 0, 0), (383999, 337, 0, 4, 2), (17, 5, 1023, 1, 0)]]
    return out


def main():
    res = bind.ref_philox_kat(cases())
    if res is None:
        raise SystemExit("oracle/_ref/philox_kat is not built (make -C oracle ref)")
    rows = [{"ctr": list(r[:4]), "key": list(r[4:6]), "out": list(r[6:])} for r in res]
    path = ROOT / "tests" / "golden" / "philox4x32_8_kat.json"
    path.write_text(json.dumps({"generator": "Random123 philox4x32_R(8, ctr, key)",
                                "source": "src/third_party/Random123/philox.h", "cases": rows},
                               indent=0))
    print(f"wrote {len(rows)} vectors to {path}")


if __name__ == "__main__":
    main()
