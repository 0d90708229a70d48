"""Write the golden fixtures under tests/golden/.

known_answers.json -- the reference's structured correctness cases
(src/main.cu:181-222) with their answers from the definition-level brute
force, plus the strict/incl answers of the tie-heavy known-answer vectors
used by tests/test_oracle.py.  Each case's expected output is also pinned
independently in tests/test_oracle.py by closed forms or hand derivation.

generators.json -- first values of the reference's input generators, so a
regression in the restated mt19937 / glibc rand() shows up without libc:
std::mt19937 seeds 42/43/100/101 (src/main.cu:246-252) and glibc rand()
after srand(1234/99/7/5/1) (src/bench_spt_le.cu:93-125, src/main.cu:411).

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def structured_cases():
    B = 128 * 4  # the reference's block size (src/main.cu:182)
    return [
        ("single element", [42]),
        ("two ascending", [1, 2]),
        ("two descending", [2, 1]),
        ("all equal", [5, 5, 5, 5, 5]),
        ("ascending", [1, 2, 3, 4, 5, 6, 7, 8]),
        ("descending", [8, 7, 6, 5, 4, 3, 2, 1]),
        ("alternating", [3, 1, 4, 1, 5, 9, 2, 6]),
        ("exact block (B=512)", list(range(B))),
        ("two full blocks", [(i % 7) * 13 for i in range(2 * B)]),
        ("non-power-of-two", [i % 17 for i in range(B + 13)]),
        # tie and sentinel vectors (strict vs incl differ here)
        ("ties", [2, 2, 1, 1, 2, 1, 3, 3]),
        ("intmax", [2147483647, 5, 2147483647, 2147483647, -2147483648, 2147483647]),
    ]


def main():
    cases = []
    for label, data in structured_cases():
        cases.append({
            "label": label,
            "input": data,
            "strict": oracle.apsep_brute(data, incl=False).tolist(),
            "incl": oracle.apsep_brute(data, incl=True).tolist(),
        })
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        f.write('{"source": "src/main.cu:181-222 (+ tie/INT_MAX vectors)",\n')
        f.write('"cases": [\n')
        f.write(",\n".join(json.dumps(c, separators=(",", ":")) for c in cases))
        f.write("\n]}\n")

    gens = {"mt19937": {}, "glibc_rand": {}}
    for seed in (42, 43, 100, 101):
        r = oracle.MT19937(seed)
        gens["mt19937"][str(seed)] = [r() for _ in range(8)]
    for seed in (1, 5, 7, 99, 1234):
        gens["glibc_rand"][str(seed)] = oracle.glibc_rand(seed, 8).tolist()
    with open(os.path.join(HERE, "generators.json"), "w") as f:
        json.dump(gens, f)
        f.write("\n")


if __name__ == "__main__":
    main()
