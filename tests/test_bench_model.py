"""The algorithmic bytes that bench.py divides by each kernel's time (kernel_bytes) against
the HBM traffic rocprofv3 measured for the same kernels at 128x129x128 (the committed
profiles/traffic_{fp64,mixed}_c1.txt: FETCH_SIZE x 2 + WRITE_SIZE per launch).  CPU only.
Every kernel must move within 0.95..1.25 of its algorithmic bytes: more means re-reads or
straddled cache lines (the FP32 Lamb rows once made the span transform fetch 1.75x its
reads), less means the model counts bytes the kernel does not move."""
import os
import re

import pytest

import rpcf
from helpers import params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KEYS = {"k_velocity_tile": "k_velocity", "k_cheb_s2p_h": "k_cheb_s2p", "k_s2p_mfma": "k_cheb_s2p",
        "k_span_inv": "k_span_inv", "k_lamb_phys": "k_lamb_phys", "k_span_fwd": "k_span_fwd",
        "k_p2s_dmfma": "k_p2s_nl", "k_p2s_mfma": "k_p2s_nl", "k_lin_tv_stream": "k_lin_tv",
        "k_lin_solve": "k_lin_solve"}


@pytest.mark.parametrize("form,precision", [("fp64", 0), ("mixed", 1)])
def test_kernel_bytes_match_measured_traffic(form, precision):
    import sys
    sys.path.insert(0, ROOT)
    import bench
    prm = params("C1")
    prm.precision = precision
    model = bench.kernel_bytes(rpcf.sizes(rpcf.load(rpcf.HIP_LIB), prm), prm)
    seen = set()
    for line in open(os.path.join(ROOT, "profiles", f"traffic_{form}_c1.txt")):
        m = re.match(r"(k_[a-z0-9_]+)\S*(?:\s*<[^>]*>)?\s+.*\s(\d+)$", line.strip())
        if not m or m.group(1) not in KEYS:
            continue
        key = KEYS[m.group(1)]
        ratio = float(m.group(2)) / model[key]
        print(form, key, "%.3f" % ratio)
        assert 0.95 < ratio < 1.25, (form, key, ratio)
        seen.add(key)
    assert seen == set(model), seen ^ set(model)
