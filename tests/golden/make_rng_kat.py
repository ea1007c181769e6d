"""Known-answer vectors for the reference's random numbers, worked out with
Python integers straight from the definitions in
crates/stitch-core/src/recon/dense.rs (not from the oracle):

- rnd (dense.rs:273-279): s ^= s << 13; s ^= s >> 7; s ^= s << 17 on u64, the
  draw is (s >> 40) as f32 / 2^24;
- the initial seed of pixel i (dense.rs:374): i * 0x9E37_79B9_7F4A_7C15 | 1,
  wrapping;
- the refinement seed (dense.rs:446-448): ((i << 8) ^ (2 it + parity)) *
  0x2545_F491_4F6C_DD1D | 1, wrapping (in Rust `<<` binds tighter than `^`).

Run from the repository root to rewrite rng_kat.json.
"""
import json
import os

M = (1 << 64) - 1


def rnd(s):
    s ^= (s << 13) & M
    s ^= s >> 7
    s ^= (s << 17) & M
    return s, (s >> 40) / float(1 << 24)  # a 24-bit integer over 2^24: exact in f32


def seed_init(i):
    return ((i * 0x9E3779B97F4A7C15) & M) | 1


def seed_refine(i, it, parity):
    return (((((i << 8) & M) ^ (2 * it + parity)) * 0x2545F4914F6CDD1D) & M) | 1


def main():
    pixels = [0, 1, 2, 3, 1 << 20, 2999999, 4319999, (1 << 32) + 5]
    out = {"seed_init": [], "seed_refine": [], "draws": []}
    for i in pixels:
        out["seed_init"].append([i, format(seed_init(i), "016x")])
        for it in (0, 3, 5):
            for parity in (0, 1):
                out["seed_refine"].append([i, it, parity, format(seed_refine(i, it, parity), "016x")])
    for s0 in [seed_init(0), seed_init(1), seed_init(1 << 20), seed_refine(7, 2, 1), 1]:
        s, draws = s0, []
        for _ in range(8):
            s, v = rnd(s)
            draws.append([format(s, "016x"), v])
        out["draws"].append({"seed": format(s0, "016x"), "draws": draws})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_kat.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
