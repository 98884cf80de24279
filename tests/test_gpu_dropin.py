"""The drop-in, proved with the reference's own code driving the library.

oracle/_ref/libanukulan_ref_hip.so is the reference's CPU path compiled from
its sources plus the LinAlgBackend adapter (oracle/hip_adapter.cpp) that
forwards every backend virtual (backend.hpp:24-215) to the C ABI of
libanukulan_hip.so. Two checks run the reference unchanged on the device:

* its solve_pdhg (pdhg.cpp:377-1285) with options.backend set to the HIP
  backend, against the same solve_pdhg on its serial CPU backend, on the
  12-period refinery plan at 1e-4 and 1e-8, and on the Wyndor LP: the same
  status, the objective within 1e-8 relative, iterations within 2%;
* its backend contract test (tests/test_backend.cpp, run_contract and
  test_backends_agree at :713-893) built with its GPU slot cuda_backend()
  bound to the adapter (oracle/_ref/test_backend_hip): every contract check
  passes on the HIP library, and the cross-check against the CPU backend
  holds to the reference's own tolerances.

Both artefacts are built by `make -C oracle ref-hip` where the reference
sources are; without them the tests skip.
"""
import json
import os
import subprocess

import pytest

from anukulan_amd import generators as G

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def relerr(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.fixture(scope="module")
def ref_hip():
    import _oracle
    lib = _oracle.load_ref_hip()
    if lib is None:
        pytest.skip("oracle/_ref/libanukulan_ref_hip.so not built (needs the reference sources)")
    return _oracle, lib


@pytest.mark.parametrize("tol", ["1e-4", "1e-8"])
def test_reference_solve_pdhg_on_the_hip_backend(ref_hip, tol):
    O, lib = ref_hip
    fx = json.load(open(os.path.join(HERE, "golden", "pdhg_reference.json")))["refinery"]["12"]
    R = O.RefLp.from_natural(G.refinery_model(12), lib)
    cpu = R.solve(threads=1, tolerance=float(tol))
    hip = R.solve(threads=-1, tolerance=float(tol))
    print(f"\nrefinery 12 at {tol}: reference solve_pdhg on the HIP backend {hip.iterations} "
          f"iterations, objective {hip.objective!r}; on its CPU backend {cpu.iterations}, "
          f"{cpu.objective!r}; fixture {fx[tol]['iterations']}")
    assert cpu.iterations == fx[tol]["iterations"]  # the adapter build is the reference's
    assert hip.status == cpu.status == "optimal"
    assert relerr(hip.objective, cpu.objective) <= 1e-8
    assert abs(hip.iterations - cpu.iterations) <= 0.02 * cpu.iterations
    assert abs(hip.matrix_norm_estimate - cpu.matrix_norm_estimate) <= 1e-12 * cpu.matrix_norm_estimate


def test_reference_solve_pdhg_on_the_hip_backend_known_lp(ref_hip):
    from golden.known_lps import KNOWN, natural
    O, lib = ref_hip
    R = O.RefLp.from_natural(natural("wyndor"), lib)
    cpu = R.solve(threads=1, tolerance=1e-8)
    hip = R.solve(threads=-1, tolerance=1e-8)
    assert hip.status == "optimal"
    assert hip.iterations == cpu.iterations
    assert abs(hip.objective - KNOWN["wyndor"]["optimum"]) <= 1e-6 * 36.0


def test_reference_backend_contract_on_the_hip_library():
    import _oracle
    exe = _oracle.CONTRACT_PATH
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/test_backend_hip not built (needs the reference sources)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print("\n" + r.stdout + r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "contract: cpu" in r.stdout
    assert "contract: hip" in r.stdout  # the GPU slot ran, on this library
    assert "cross-check: cpu vs hip" in r.stdout
    assert "ok   test_backend" in r.stdout
    assert "FAIL" not in r.stdout + r.stderr
