"""Gauge-configuration files: the library's reader and writer
(lym_config_read / lym_config_write, the reference's
ReadStartGaugeConfiguration.cpp:37-500 and OutputSweep.cpp:19-440) against
the oracle restatement in oracle/gauge_io.py.

The oracle itself is pinned first: its records equal what Python's xdrlib
(an independent XDR implementation) packs for the same numbers, its
plaquette meets closed forms, and its readers invert its writers.  Then the
library must write the same bytes as the oracle and read the oracle's files
to the same links.  CPU only (host code)."""
import os
import sys
import xdrlib

import numpy as np
import pytest

import leonardym_amd as lym

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import gauge_io as orc  # noqa: E402

DIMS = (4, 2, 2, 6)
V = int(np.prod(DIMS))


def su3(seed):
    return lym.gen_su3_links(DIMS, 0.6, seed)


def su2(seed):
    return lym.gen_su2_links(DIMS, 0.6, seed)


# ------------------------------------------------------------------ oracle pinned
def test_oracle_leonard_record_is_xdr():
    U = su3(1)
    C = orc.lex_coords(DIMS)
    s = 37
    p = xdrlib.Packer()
    for v in C[s]:
        p.pack_int(int(v))
    for mu in range(4):
        for i in range(3):
            for j in range(3):
                p.pack_double(U[s, mu, i, j].real)
                p.pack_double(U[s, mu, i, j].imag)
    assert orc.leonard_record_bytes(C[s], U[s]) == p.get_buffer()


@pytest.mark.parametrize("N", [2, 3])
def test_oracle_muenster_is_xdr(N, tmp_path):
    U = su2(2) if N == 2 else su3(2)
    path = str(tmp_path / "cfg")
    orc.write_muenster(path, DIMS, U, beta=2.25, kappa=0.1775, sweep=41, plaquette=0.625)
    p = xdrlib.Packer()
    for t in range(DIMS[3]):
        for z in range(DIMS[2]):
            for y in range(DIMS[1]):
                for x in range(DIMS[0]):
                    s = ((x * DIMS[1] + y) * DIMS[2] + z) * DIMS[3] + t
                    for mu in range(4):
                        W = U[s, mu]
                        if N == 2:
                            D = np.conj(W.T)  # htrans: the file holds its first row
                            for jj in range(2):
                                p.pack_float(D[0, jj].real)
                                p.pack_float(D[0, jj].imag)
                        else:
                            for ii in range(3):
                                for jj in range(2):
                                    p.pack_float(W[ii, jj].real)
                                    p.pack_float(-W[ii, jj].imag)
    if N == 2:
        for v in (0.625, 0.0, 0.0, 2.25, 0.1775, 0.0, 0.0, 0.0):
            p.pack_float(v)
        for v in [41, 0, 0, 0, 0, 0] + list(DIMS) + [0]:
            p.pack_int(v)
    else:
        for v in DIMS:
            p.pack_int(v)
        p.pack_double(float(np.float32(2.25)))
        p.pack_double(-1.0 / 12.0)
        p.pack_double(0.625)
    assert open(path, "rb").read() == p.get_buffer()


def test_oracle_plaquette_closed_forms():
    # cold links: every plaquette is the identity
    cold = np.broadcast_to(np.eye(3, dtype=complex), (V, 4, 3, 3)).copy()
    assert orc.temporal_plaquette(DIMS, cold) == pytest.approx(1.0, abs=1e-15)
    # abelian field U_t(x) = exp(i phi x T), T = diag(1, -1, 0), other links 1:
    # the (x, t) plaquettes are exp(i phi T) except at x = Lx - 1, where the
    # neighbour wraps and they are exp(-i phi (Lx - 1) T); (y, t), (z, t) are 1
    phi = 0.37
    U = cold.copy()
    C = orc.lex_coords(DIMS)
    for s in range(V):
        a = phi * C[s, 0]
        U[s, 3] = np.diag([np.exp(1j * a), np.exp(-1j * a), 1.0])
    Lx = DIMS[0]
    per_site_x = ((Lx - 1) * (2 * np.cos(phi) + 1) + (2 * np.cos(phi * (Lx - 1)) + 1)) / Lx
    want = (per_site_x + 3 + 3) / (3 * 3)
    assert orc.temporal_plaquette(DIMS, U) == pytest.approx(want, abs=1e-14)
    # a plausible slip -- U_t(x) for U_t(x + mu) -- changes the value
    assert abs(want - (3 + 3 + 3) / 9) > 1e-3


@pytest.mark.parametrize("nproc", [1, 2])
def test_oracle_leonard_round_trip(nproc, tmp_path):
    U = su3(3)
    orc.write_leonard(str(tmp_path) + "/", "conf", 7, DIMS, U, nproc=nproc)
    R, plaq = orc.read_leonard(str(tmp_path) + "/", "conf", 7, DIMS, 3)
    assert np.array_equal(R, U)
    assert plaq == pytest.approx(orc.temporal_plaquette(DIMS, U), abs=1e-12)


@pytest.mark.parametrize("N", [2, 3])
def test_oracle_muenster_round_trip(N, tmp_path):
    U = su2(4) if N == 2 else su3(4)
    path = orc.muenster_path(str(tmp_path) + "/", "run", 3, N)
    orc.write_muenster(path, DIMS, U, beta=5.4, kappa=0.135, sweep=9)
    R, info = orc.read_muenster(path, DIMS, N)
    # float storage: the links come back to float rounding
    assert np.max(np.abs(R - U)) < 3e-7
    eye = np.eye(N)
    UU = np.max(np.abs(R @ np.conj(np.swapaxes(R, -1, -2)) - eye))
    if N == 2:
        # rebuilt from one row: SU(2) form exactly, unitary to float rounding (not re-unitarised)
        assert np.array_equal(R[..., 1, 1], np.conj(R[..., 0, 0]))
        assert np.array_equal(R[..., 0, 1], -np.conj(R[..., 1, 0]))
        assert UU < 1e-6
    else:
        # SU(3): re-unitarised from two columns
        assert UU < 1e-14
    if N == 3:
        assert np.max(np.abs(np.linalg.det(R) - 1)) < 1e-14
        assert info["c1"] == -1.0 / 12.0
    else:
        assert info["kappa"] == pytest.approx(0.135, rel=1e-7) and info["sweep"] == 9
    assert info["beta"] == pytest.approx(5.4, rel=1e-7)


# ------------------------------------------------------------------ library vs oracle
def test_library_plaquette():
    for U in (su3(5), su2(5)):
        assert lym.temporal_plaquette(DIMS, U) == pytest.approx(orc.temporal_plaquette(DIMS, U), abs=1e-14)


def test_library_writes_oracle_bytes_leonard(tmp_path):
    U = su3(6)
    a, b = tmp_path / "lib", tmp_path / "orc"
    a.mkdir()
    b.mkdir()
    info = lym.config_write(lym.FORMAT_LEONARD, str(a) + "/", "cfg", 12, DIMS, U)
    orc.write_leonard(str(b) + "/", "cfg", 12, DIMS, U, plaquette=info.plaquette)
    assert (a / "cfg_12_0.txt").read_bytes() == (b / "cfg_12_0.txt").read_bytes()
    da = (a / "cfg_12.descriptor.txt").read_text().split()
    db = (b / "cfg_12.descriptor.txt").read_text().split()
    assert da[:4] == db[:4] and da[8:10] == db[8:10]
    assert float(da[10]) == pytest.approx(orc.temporal_plaquette(DIMS, U), abs=1e-12)


@pytest.mark.parametrize("N", [2, 3])
def test_library_writes_oracle_bytes_muenster(N, tmp_path):
    U = su2(7) if N == 2 else su3(7)
    a, b = tmp_path / "lib", tmp_path / "orc"
    a.mkdir()
    b.mkdir()
    info = lym.config_write(lym.FORMAT_MUENSTER, str(a) + "/", "ens", 4, DIMS, U, beta=2.2, kappa=0.18, sweep=4)
    orc.write_muenster(orc.muenster_path(str(b) + "/", "ens", 4, N), DIMS, U, beta=2.2, kappa=0.18, sweep=4,
                       plaquette=info.plaquette)
    fa = orc.muenster_path(str(a) + "/", "ens", 4, N)
    fb = orc.muenster_path(str(b) + "/", "ens", 4, N)
    assert open(fa, "rb").read() == open(fb, "rb").read()


@pytest.mark.parametrize("nproc", [1, 2, 4])
def test_library_reads_oracle_leonard(nproc, tmp_path):
    U = su3(8)
    orc.write_leonard(str(tmp_path) + "/", "c", 0, DIMS, U, nproc=nproc)
    R, info = lym.config_read(lym.FORMAT_LEONARD, str(tmp_path) + "/", "c", 0, DIMS, 3)
    assert np.array_equal(R, U)
    assert info.plaquette == pytest.approx(orc.temporal_plaquette(DIMS, U), abs=1e-12)


@pytest.mark.parametrize("N", [2, 3])
def test_library_reads_oracle_muenster(N, tmp_path):
    U = su2(9) if N == 2 else su3(9)
    path = orc.muenster_path(str(tmp_path) + "/", "e", 2, N)
    orc.write_muenster(path, DIMS, U, beta=2.3, kappa=0.16, sweep=77)
    R, info = lym.config_read(lym.FORMAT_MUENSTER, str(tmp_path) + "/", "e", 2, DIMS, N)
    Ro, io = orc.read_muenster(path, DIMS, N)
    assert np.max(np.abs(R - Ro)) < 1e-15
    assert info.beta == pytest.approx(io["beta"], abs=0) and info.plaquette == pytest.approx(io["plaquette"], abs=0)
    if N == 2:
        assert info.kappa == pytest.approx(io["kappa"], abs=0) and info.sweep == 77
    else:
        assert info.c1 == -1.0 / 12.0


def test_library_round_trip_and_adjoint(tmp_path):
    # an SU(2) ensemble read back feeds the adjoint fermion links
    U = su2(10)
    lym.config_write(lym.FORMAT_LEONARD, str(tmp_path) + "/", "a", 1, DIMS, U)
    R, _ = lym.config_read(lym.FORMAT_LEONARD, str(tmp_path) + "/", "a", 1, DIMS, 2)
    assert np.array_equal(R, U)
    assert np.array_equal(lym.su2_to_adjoint(R), lym.su2_to_adjoint(U))


def test_library_errors(tmp_path):
    U = su3(11)
    d = str(tmp_path) + "/"
    lym.config_write(lym.FORMAT_LEONARD, d, "x", 1, DIMS, U)
    with pytest.raises(lym.LymError, match="different lattice size"):
        lym.config_read(lym.FORMAT_LEONARD, d, "x", 1, (4, 2, 2, 4), 3)
    with pytest.raises(lym.LymError, match="cannot open"):
        lym.config_read(lym.FORMAT_LEONARD, d, "x", 2, DIMS, 3)
    # a truncated process file is an error (the reference keeps stale links)
    f = tmp_path / "x_1_0.txt"
    f.write_bytes(f.read_bytes()[:-100])
    with pytest.raises(lym.LymError, match="ends after"):
        lym.config_read(lym.FORMAT_LEONARD, d, "x", 1, DIMS, 3)
    lym.config_write(lym.FORMAT_MUENSTER, d, "m", 1, DIMS, U, beta=6.0)
    with pytest.raises(lym.LymError, match="lattice size is different"):
        lym.config_read(lym.FORMAT_MUENSTER, d, "m", 1, (2, 4, 2, 6), 3)
    g = tmp_path / "m.config-1"
    g.write_bytes(g.read_bytes()[:1000])
    with pytest.raises(lym.LymError, match="too short"):
        lym.config_read(lym.FORMAT_MUENSTER, d, "m", 1, DIMS, 3)
    with pytest.raises(lym.LymError):
        lym.config_write(7, d, "z", 1, DIMS, U)
