"""Writes sobol_golden.json: Owen-scrambled Sobol integers from the oracle at the
parity-critical indices (index 0, config-maximum indices, random u32) and all 256
dimensions. The oracle is pinned independently by scipy's unscrambled Sobol and the
stratification test; this file freezes the result as a regression fixture."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle  # noqa: E402

rng = np.random.default_rng(0x50B01)
idx = np.concatenate([[0, 1, 2, 3, 4, 3_932_100, 522_546_948, 264_240_900, 114_278_276, 0xFFFFFFFE, 0xFFFFFFFF],
                      rng.integers(0, 2 ** 32, 53, dtype=np.uint64)]).astype(np.uint32)
dims = np.arange(256, dtype=np.uint32)
I, D = np.meshgrid(idx, dims, indexing="ij")
bits = pyoracle.sobol_bits(I.ravel(), D.ravel())
json.dump({"index": I.ravel().tolist(), "dim": D.ravel().tolist(), "bits": bits.tolist()}, open(os.path.join(HERE, "sobol_golden.json"), "w"))
print("wrote", bits.size, "vectors")
