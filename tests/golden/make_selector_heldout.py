"""Writes the selector's held-out fixture: features and measured variant runtimes of the 512
held-out graphs of the mixed-batch distribution (seeds 2,000,000 .. 2,000,511; the training
graphs use seeds 1,000,000 ..).  Needs a GPU (the labels are the library's own timings).

    python tests/golden/make_selector_heldout.py OUT.npz

OUT.npz holds F float32 [512, 16] (pbf_batch_features), T float32 [512, 9] (log2 ms of each
variant's per-graph pipeline, parbutterfly_amd.selector.measure) and seeds int64 [512].  The
committed copy is tests/golden/selector_heldout512.npz; profiles/selector_heldout512.json is
the selector's evaluation on it (python -m parbutterfly_amd.selector evaluate-fixture).
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

FIRST_SEED, COUNT = 2_000_000, 512


def main():
    import parbutterfly_amd as pb
    from parbutterfly_amd import selector as S
    out = sys.argv[1]
    ctx = pb.Context(0)
    with tempfile.TemporaryDirectory() as td:
        S.collect(ctx, FIRST_SEED, COUNT, td, chunk=COUNT)
        F, T = S.load_dir(td)
    np.savez(out, F=F.astype(np.float32), T=T.astype(np.float32),
             seeds=np.arange(FIRST_SEED, FIRST_SEED + COUNT, dtype=np.int64))
    print(f"wrote {out}: F {F.shape}, T {T.shape}", flush=True)


if __name__ == "__main__":
    main()
