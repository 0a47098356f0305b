"""The serving loop's scheduling decisions on a mock device, on the CPU:
admission in registration order and one at a time, deferral when an SCF does
not fit beside the resident ones, eviction of the least recently used idle
SCF (never a busy or queued one, and only when it makes room), the oldest
runnable request first, failure of an SCF larger than the budget, release of
a closed SCF.  tests/native/server_policy.cpp drives the same policy
functions the server uses (occ_amd/csrc/server_policy.hpp); it is built with
AddressSanitizer and UndefinedBehaviorSanitizer."""
import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parent.parent


def test_server_policy_on_a_mock_device(tmp_path):
    exe = tmp_path / "server_policy"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-fsanitize=address,undefined",
                    "-fno-omit-frame-pointer", f"-I{ROOT / 'occ_amd' / 'csrc'}",
                    str(ROOT / "tests" / "native" / "server_policy.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 failed checks" in out.stdout
