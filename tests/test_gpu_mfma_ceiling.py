"""GPU: the sustained MFMA rate of the device (tests/native/mfma_ceiling.hip,
built by __graft_entry__.build()).  It calibrates the contraction rooflines:
FP64 MFMA reaches its published dense peak (78.6 TFLOP/s) only with the
accumulators in ArchVGPRs (launch bound of two waves per SIMD); with them in
AccVGPRs it sustains about half.  FP32 MFMA comes close to its 157.3 either
way.  The test prints the rates and checks their ordering."""
import pathlib
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = pathlib.Path(__file__).resolve().parent / "native" / "mfma_ceiling"


def test_mfma_ceiling_rates():
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    out = subprocess.run([str(EXE)], check=True, capture_output=True, text=True,
                         timeout=120).stdout
    print(out)
    rates = {}
    for line in out.splitlines():
        m = re.match(r"(fp\d+) accumulators (\d+) bound (\d) workgroups (\d+): .* ([\d.]+) "
                     r"TFLOP/s", line)
        assert m, line
        rates[(m.group(1), int(m.group(2)), int(m.group(3)), int(m.group(4)))] = float(m.group(5))
    assert len(rates) == 18 and all(v > 0 for v in rates.values())
    best64 = max(v for k, v in rates.items() if k[0] == "fp64")
    best32 = max(v for k, v in rates.items() if k[0] == "fp32")
    assert best32 > best64
    # FP64: accumulators in ArchVGPRs (bound 2) beat AccVGPRs (bound 1)
    assert rates["fp64", 16, 2, 4096] > rates["fp64", 16, 1, 4096]
