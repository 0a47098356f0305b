"""GPU: two device rates the contraction kernels are shaped by
(tests/native/lds_atomic_rates.hip, built by __graft_entry__.build()).

- An MFMA operand read (ds_read_b128, lane l: row l % 16, 4-float group
  l / 16) from 64-byte rows whose 16-byte slots are XOR-swizzled by
  (row / 2) % 4 runs at the LDS's full rate; the padded row strides (144 or
  80 bytes) and the other layouts take about twice as long.  k_vxc_gemm_f32
  stores its operands that way.
- FP64 atomic adds in the Vxc scatter's shape run several times slower than
  plain stores of the same shape, and at the same rate into a matrix that
  fits the L2 as into one that does not: the scatter is limited by the
  atomics' own rate, so fewer atomics (the folded diagonal blocks) is the
  lever, not locality.
The test prints the rates and checks those orderings."""
import pathlib
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = pathlib.Path(__file__).resolve().parent / "native" / "lds_atomic_rates"


def test_lds_and_atomic_rates():
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    out = subprocess.run([str(EXE)], check=True, capture_output=True, text=True,
                         timeout=120).stdout
    print(out)
    lds, scat = {}, {}
    for line in out.splitlines():
        m = re.match(r"lds layout (\d) \(.*\): .* ([\d.]+) cycles", line)
        if m:
            lds[int(m.group(1))] = float(m.group(2))
            continue
        m = re.match(r"scatter rows (\d+) .*: atomics ([\d.e+]+) per s, stores ([\d.e+]+) per s",
                     line)
        assert m, line
        scat[int(m.group(1))] = (float(m.group(2)), float(m.group(3)))
    assert len(lds) == 6 and len(scat) == 2
    # the swizzled 64-byte rows at the linear layout's rate, faster than padding
    assert lds[1] < 1.25 * lds[0]
    assert lds[4] > 1.3 * lds[1] and lds[5] > 1.3 * lds[1]
    for atomics, stores in scat.values():
        assert 0 < atomics < stores
    big, small = scat[3072][0], scat[96][0]
    assert 0.5 < big / small < 2.0
