"""The product's HIP kernel sources, compiled for the CPU by tests/emu (a
workgroup's lanes run as fibers between barriers, no device code), agree with the
oracle on small cases.  Runs in a subprocess so the emulation library never
replaces libcns.so in this test process."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")


def test_kernel_sources_on_cpu_emulation():
    subprocess.run(["make", "-s", "-j8", "-C", EMU, "_build/libcns_emu.so"], check=True)
    lib = os.path.join(EMU, "_build", "libcns_emu.so")
    env = dict(os.environ, OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, os.path.join(EMU, "run_emu_parity.py"), lib], capture_output=True,
                       text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stdout + r.stderr
