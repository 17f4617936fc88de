"""Writes the committed fixtures of tests/golden from the reference checkout (run here,
where /root/reference exists; the tests only read the outputs).

  ParDiskANN_Ls_75.bin  a neighbour file the reference wrote (cuda/comparacion1.cu:126-138):
                        29,316 x 10 ids on MSSPACEV-1M, 0xFFFFFFFF for missing neighbours.
                        Used as a format fixture (the base vectors and graph are absent).
  hash_vectors.json     hash64_2 / hash64 known answers.  The multiply constants, xor-shift
                        amounts and their order are parsed from the text of
                        cuda/utiles_gpu.cuh:12-31 and evaluated with Python integers, so the
                        vectors come from the reference's definition, not from the oracle.
"""
import json
import os
import re
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1


def parse_function(text, name):
    body = re.search(name + r"\(uint64_t \w+\) \{(.*?)\n\}", text, re.S).group(1)
    return [ln.strip() for ln in body.strip().splitlines() if ln.strip() and not ln.strip().startswith("//")]


def evaluate(lines, x):
    """Interprets the handful of statement shapes used by the two hash functions."""
    env = {}
    var = None
    for ln in lines:
        ln = ln.rstrip(";")
        m = re.match(r"uint64_t (\w+) = (\w+) \* (\d+)ull \+ (\d+)ull", ln)
        if m:
            var = m.group(1)
            env[var] = (x * int(m.group(3)) + int(m.group(4))) & M64
            continue
        m = re.match(r"(\w+) = \((\w+) \^ \((\w+) >> (\d+)\)\) \* (0x[0-9a-f]+)ull", ln)
        if m:
            var = m.group(1)
            v = env.get(var, x)
            env[var] = ((v ^ (v >> int(m.group(4)))) * int(m.group(5), 16)) & M64
            continue
        m = re.match(r"(\w+) = (\w+) \^ \((\w+) >> (\d+)\)", ln)
        if m:
            var = m.group(1)
            v = env.get(var, x)
            env[var] = (v ^ (v >> int(m.group(4)))) & M64
            continue
        m = re.match(r"(\w+) \^= (\w+) (<<|>>) (\d+)", ln)
        if m:
            var = m.group(1)
            v = env[var]
            s = int(m.group(4))
            env[var] = (v ^ ((v << s) if m.group(3) == "<<" else (v >> s))) & M64
            continue
        m = re.match(r"(\w+) \*= (\d+)ull", ln)
        if m:
            var = m.group(1)
            env[var] = (env[var] * int(m.group(2))) & M64
            continue
        m = re.match(r"return (\w+)", ln)
        if m:
            return env.get(m.group(1), x)
        raise ValueError("unparsed statement: " + ln)
    raise ValueError("no return")


def main():
    text = open(os.path.join(REF, "cuda/utiles_gpu.cuh")).read()
    h2 = parse_function(text, "hash64_2")
    h1 = parse_function(text, "hash64")
    xs = [0, 1, 2, 3, 7, 31, 1000, 999999, 123456789, 0xFFFFFFFE, 0xFFFFFFFF, 1 << 40,
          0x9E3779B97F4A7C15, M64]
    out = {"inputs": [str(x) for x in xs],
           "hash64_2": [str(evaluate(h2, x)) for x in xs],
           "hash64": [str(evaluate(h1, x)) for x in xs],
           "source": "cuda/utiles_gpu.cuh:12-31"}
    with open(os.path.join(HERE, "hash_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    shutil.copyfile(os.path.join(REF, "resultados i7/gpu_t24/vecinos/ParDiskANN_Ls_75.bin"),
                    os.path.join(HERE, "ParDiskANN_Ls_75.bin"))


if __name__ == "__main__":
    main()
