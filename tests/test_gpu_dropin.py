"""The C++ drop-in (leonardym_amd/host) used the way the reference's callers
use DiracOperator / PreconditionedBiCGStab / ConjugateGradient, checked
against the oracle.  The driver binary is built by __graft_entry__.build()."""
import os
import subprocess

import numpy as np
import pytest

from helpers import GAMMA5, K_CLOVER, OracleField, fold_bc, rel_err_per_site

import leonardym_amd as lym

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(lym.__file__), "host", "dropin_check")


def test_dropin_driver_matches_oracle(tmp_path):
    assert os.path.exists(DRIVER), "build the drop-in driver first (__graft_entry__.build())"
    dims = (8, 8, 8, 8)
    V = int(np.prod(dims))
    kappa, csw = 0.125, 1.0
    U = fold_bc(dims, lym.gen_su3_links(dims, 0.5, 77))
    links = tmp_path / "links.bin"
    np.ascontiguousarray(np.swapaxes(U, -1, -2)).tofile(links)
    out = tmp_path / "out.bin"
    print("running", DRIVER, flush=True)
    r = subprocess.run([DRIVER, *map(str, dims), repr(kappa), repr(csw), str(links), str(out)],
                       capture_output=True, text=True, timeout=300)
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    assert "expected error" in r.stderr
    res = np.fromfile(out, dtype=np.complex128).reshape(-1, V, 4, 3)
    assert res.shape[0] == 18
    orc = OracleField(dims, U, kappa=kappa, csw=csw, need_clover=True)
    # 12 propagator columns: D x = e_(alpha, c) at the origin
    for k in range(12):
        e = np.zeros((V, 4, 3), complex)
        e[0, k // 3, k % 3] = 1
        Dx = orc.apply(K_CLOVER, res[k], gamma5=False)
        assert np.linalg.norm(Dx - e) <= 1e-10
    psi = lym.gaussian_spinor(V, 11)
    psi2 = lym.gaussian_spinor(V, 12)
    assert rel_err_per_site(res[12], orc.apply(K_CLOVER, psi)) < 1e-13
    assert rel_err_per_site(res[13], orc.apply(K_CLOVER, psi, gamma5=False)) < 1e-13
    assert rel_err_per_site(res[14], orc.apply(K_CLOVER, psi, v2=psi2, alpha=0.25 - 0.5j)) < 1e-13
    assert rel_err_per_site(res[15], orc.apply(K_CLOVER, psi, power=2)) < 1e-12
    Q2x = orc.apply(K_CLOVER, res[16], power=2)
    assert np.linalg.norm(Q2x - psi) / np.linalg.norm(psi) <= 1e-10
    # getSquareRoot(getSquare(Q)) applies Q again
    assert rel_err_per_site(res[17], orc.apply(K_CLOVER, psi)) < 1e-13
