"""Writes sobol_ref_golden.npz: Owen-scrambled Sobol vectors computed by the reference's
own Python generator, scripts/generate_luts_sobol.py.

Run in a container that holds the reference tree (it is read from /root/reference and
executed with runpy inside a temporary directory, which receives the header and object
file the script writes; nothing is written to the reference tree):

    python tests/golden/make_sobol_ref_golden.py

The script's functions give the 32-bit results of math_random.h:48-85 on u32 inputs:
  - compute_V (generate_luts_sobol.py:90-95) builds the direction numbers the C++ header
    is generated from (math/CMakeLists.txt:71-85 runs the script with -d 256 -s 32);
  - gray_code, hash, reverse_bits and owen_scramble (:100-125) restate
    SobolGenerator::{gray_code, hash, reverse_bits, owen_scramble}; Python integers are
    unbounded, but only *, ^, + and | touch x before reverse_bits masks it to 32 bits,
    and the low 32 bits of those depend only on the low 32 bits of the operands;
  - sobol_sample (:128-138) returns accum / 2^32 in double (0 at index 0), whose float32
    rounding equals the C++ float(accum) * 2^-32 for every nonzero index.

The fixture holds:
  table  uint32 [256, 32]  the direction table parsed from the header the script writes
                           (row 255 is the zero row of the [256][32] declaration)
  index  uint32 [N]        index 0, the configuration-maximum indices, edge values and
                           random u32 indices
  bits   uint32 [N, 256]   owen_scramble(XOR of V[dim][i] over the set bits of
                           gray_code(index)), seed 0xDEADBEEF
  value  float32 [N, 256]  float32(sobol_sample(row, index)) for dims 0..254 (the script
                           has no row for dim 255); index 0 rows are left at 0 as the
                           script returns
"""
import os
import re
import runpy
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SCRIPT = "/root/reference/scripts/generate_luts_sobol.py"
JOE_KUO = "/root/reference/scripts/joe_kuo.txt"
SEED = 0xDEADBEEF


def parse_header(text):
    body = text[text.index("{") + 1:]
    rows = re.findall(r"\{([^{}]*)\}", body)
    table = np.zeros((256, 32), dtype=np.uint32)
    for r, row in enumerate(rows):
        vals = [int(v, 16) for v in row.split(",")]
        assert len(vals) == 32
        table[r] = vals
    return table, len(rows)


def main():
    cwd = os.getcwd()
    argv = sys.argv
    with tempfile.TemporaryDirectory() as tmp:
        os.chdir(tmp)
        sys.argv = [REF_SCRIPT, "-d", "256", "-s", "32", "-o", "sobol_direction_vectors", "-i", JOE_KUO]
        try:
            g = runpy.run_path(REF_SCRIPT, run_name="__main__")
            header = open(os.path.join(tmp, "sobol_direction_vectors.h")).read()
        finally:
            sys.argv = argv
            os.chdir(cwd)
    table, nrows = parse_header(header)
    rows = g["list_rows"]
    assert nrows == len(rows) == 255, (nrows, len(rows))
    V = np.zeros((256, 32), dtype=np.uint64)
    for d, row in enumerate(rows):
        V[d] = g["compute_V"](row[1], row[2], list(row[3]), 32)
    assert (V.astype(np.uint32) == table).all()

    rng = np.random.default_rng(0x50B02)
    config_max = [3_932_100, 522_546_948, 264_240_900, 114_278_276]  # 4*p*f maxima of the five workloads
    edges = [0, 1, 2, 3, 4, 5, 63, 64, 65, 1 << 16, (1 << 31) - 1, 1 << 31, 0xFFFFFFFE, 0xFFFFFFFF]
    index = np.concatenate([edges, config_max, rng.integers(0, 1 << 32, 46, dtype=np.uint64)]).astype(np.uint64)
    gray, owen = g["gray_code"], g["owen_scramble"]
    bits = np.zeros((len(index), 256), dtype=np.uint32)
    value = np.zeros((len(index), 256), dtype=np.float32)
    for n, i in enumerate(index.tolist()):
        gc = gray(i)
        for d in range(256):
            acc = 0
            for b in range(32):
                if (gc >> b) & 1:
                    acc ^= int(V[d, b])
            bits[n, d] = owen(acc, SEED) & 0xFFFFFFFF
            if d < 255 and i != 0:
                row = rows[d]  # sobol_sample extends row[3] in place, as compute_sobol_sequence avoids with deepcopy
                value[n, d] = np.float32(g["sobol_sample"]([row[0], row[1], row[2], list(row[3])], i, 32, SEED))
    out = os.path.join(HERE, "sobol_ref_golden.npz")
    np.savez_compressed(out, table=table, index=index.astype(np.uint32), bits=bits, value=value)
    print(f"wrote {out}: {bits.size} integer and {int((value != 0).sum())} float vectors")


if __name__ == "__main__":
    main()
