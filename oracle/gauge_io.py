"""CPU restatement of LeonardYM's gauge-configuration readers and writers.

TEST INFRASTRUCTURE ONLY.  This module is the oracle for the configuration
file formats: it is imported only by tests/, never by the product path
(leonardym_amd/), which has its own C++ reader and writer.

Links are numpy complex128 arrays [V, 4, N, N] with element (i, j) at
[..., i, j] (the reference's GaugeGroup::at(i, j)), sites in the t-fastest
lexicographic order idx = ((x*Ly + y)*Lz + z)*Lt + t
(source/MPILattice/LocalLayout.h:149-155).  XDR numbers are big-endian
(RFC 4506): numpy dtypes '>i4', '>f4', '>f8'.

Pinned in tests/test_gauge_io.py against Python's xdrlib (an independent XDR
implementation) for every record type, by exact round trips of the double
format, by the SU(2)/SU(3) reconstruction identities (an SU(N) link written
and read back is itself to float rounding) and by closed forms of the
temporal plaquette.
"""
import os

import numpy as np


def lex_coords(dims):
    """[V, 4] coordinates (x, y, z, t) of the lexicographic sites."""
    L = dims
    x, y, z, t = np.meshgrid(*(np.arange(n) for n in L), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel(), t.ravel()], axis=1)


def _shift(U_mu, dims, mu):
    """field at x + mu (source/MPILattice/LocalLayout.h sup tables, periodic)."""
    a = U_mu.reshape(tuple(dims) + U_mu.shape[1:])
    return np.roll(a, -1, axis=mu).reshape(U_mu.shape)


def temporal_plaquette(dims, U):
    """Plaquette::temporalPlaquette (source/wilson_loops/Plaquette.cpp:49-66):
    sum_x sum_{mu<3} Re tr[U_mu(x) U_t(x+mu) U_mu(x+t)^+ U_t(x)^+] / (3 N V)."""
    V, _, N, _ = U.shape
    tot = 0.0
    Ut = U[:, 3]
    for mu in range(3):
        Um = U[:, mu]
        P = Um @ _shift(Ut, dims, mu) @ np.conj(np.swapaxes(_shift(Um, dims, 3), -1, -2)) @ np.conj(
            np.swapaxes(Ut, -1, -2))
        tot += np.trace(P, axis1=-2, axis2=-1).real.sum()
    return tot / (3.0 * N * V)


# ------------------------------------------------------------------ leonard_format
def leonard_record_bytes(coords, U_site):
    """One record of OutputSweep.cpp:50-66: ints x, y, z, t then, for mu, i, j,
    doubles re, im of U_mu.at(i, j)."""
    head = np.asarray(coords, dtype=">i4").tobytes()
    body = np.stack([U_site.real, U_site.imag], axis=-1).astype(">f8").tobytes()  # [mu][i][j][re,im]
    return head + body


def write_leonard(directory, name, number, dims, U, nproc=1, plaquette=None):
    """OutputSweep.cpp:32-84.  With nproc > 1 the sites are dealt to the
    process files in contiguous lexicographic ranges (as an MPI run would
    leave them, each file holding its own sites with their coordinates)."""
    V = U.shape[0]
    C = lex_coords(dims)
    base = os.path.join(directory, "%s_%d" % (name, number))
    chunks = np.array_split(np.arange(V), nproc)
    for p, sites in enumerate(chunks):
        with open("%s_%d.txt" % (base, p), "wb") as f:
            for s in sites:
                f.write(leonard_record_bytes(C[s], U[s]))
    plaq = temporal_plaquette(dims, U) if plaquette is None else plaquette
    local = max(len(c) for c in chunks)
    with open(base + ".descriptor.txt", "w") as f:
        f.write("%d %d %d %d\n" % tuple(dims))
        f.write("%d 1 1 1\n" % nproc)
        f.write("%d\n%d\n" % (nproc, local))
        f.write("%.13g\n" % plaq)


def read_leonard(directory, name, number, dims, N):
    """ReadStartGaugeConfiguration.cpp:51-130."""
    base = os.path.join(directory, "%s_%d" % (name, number))
    with open(base + ".descriptor.txt") as f:
        tok = f.read().split()
    g = [int(v) for v in tok[0:4]]
    if g != list(dims):
        raise ValueError("different lattice size in reading configuration")
    nproc, local, plaq = int(tok[8]), int(tok[9]), float(tok[10])
    V = int(np.prod(dims))
    U = np.zeros((V, 4, N, N), complex)
    rec = 16 + 4 * N * N * 16
    for p in range(nproc):
        raw = open("%s_%d.txt" % (base, p), "rb").read()
        if len(raw) < local * rec:
            raise ValueError("short file")
        for r in range(local):
            b = raw[r * rec:(r + 1) * rec]
            x, y, z, t = np.frombuffer(b[:16], ">i4")
            v = np.frombuffer(b[16:], ">f8").reshape(4, N, N, 2)
            s = ((x * dims[1] + y) * dims[2] + z) * dims[3] + t
            U[s] = v[..., 0] + 1j * v[..., 1]
    return U, plaq


# ------------------------------------------------------------------ muenster_format
def muenster_site_order(dims):
    """Sites in file order: x fastest, then y, z, t (OutputSweep.cpp:117-120)."""
    Lx, Ly, Lz, Lt = dims
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx), indexing="ij")
    return (((x * Ly + y) * Lz + z) * Lt + t).ravel()


def muenster_path(directory, name, number, N):
    if N == 2:
        return os.path.join(directory, "0:%d:%s" % (number, name))
    return os.path.join(directory, "%s.config-%d" % (name, number))


def write_muenster(path, dims, U, beta=0.0, kappa=0.0, sweep=0, plaquette=None):
    """OutputSweep.cpp:87-283 (SU(2)) and :284-440 (SU(3))."""
    V, _, N, _ = U.shape
    order = muenster_site_order(dims)
    Us = U[order]
    plaq = temporal_plaquette(dims, U) if plaquette is None else plaquette
    if N == 2:
        # U^+ = htrans(U); its first row (tmp(0,0), tmp(0,1)) = conj(U00), conj(U10)
        row = np.conj(Us[:, :, :, 0])  # [V, 4, 2]: conj(U_{j0}) for j = 0, 1
        body = np.stack([row.real, row.imag], axis=-1).astype(">f4").tobytes()
        tr = np.array([plaq, 0.0, 0.0, beta, kappa, 0.0, 0.0, 0.0], ">f4").tobytes()
        tr += np.array([sweep, 0, 0, 0, 0, 0] + list(dims) + [0], ">i4").tobytes()
    else:
        cols = Us[:, :, :, 0:2]  # [V, 4, 3 rows, 2 cols]
        body = np.stack([cols.real, -cols.imag], axis=-1).astype(">f4").tobytes()
        tr = np.array(list(dims), ">i4").tobytes()
        tr += np.array([np.float32(beta), -1.0 / 12.0, plaq], ">f8").tobytes()
    with open(path, "wb") as f:
        f.write(body + tr)


def su3_reunitarise(m):
    """ReadStartGaugeConfiguration.cpp:338-366 on [..., 3, 2] columns."""
    c0 = m[..., 0]
    c0 = c0 / np.sqrt(np.sum(np.abs(c0) ** 2, axis=-1, keepdims=True))
    d = np.sum(m[..., 1] * np.conj(c0), axis=-1, keepdims=True)
    c1 = m[..., 1] - d * c0
    c1 = c1 / np.sqrt(np.sum(np.abs(c1) ** 2, axis=-1, keepdims=True))
    c2 = np.conj(np.cross(c0, c1))
    return np.stack([c0, c1, c2], axis=-1)


def read_muenster(path, dims, N):
    """ReadStartGaugeConfiguration.cpp:132-297 (SU(2)) and :299-480 (SU(3));
    returns links and the trailer as a dict."""
    raw = open(path, "rb").read()
    V = int(np.prod(dims))
    order = muenster_site_order(dims)
    U = np.zeros((V, 4, N, N), complex)
    if N == 2:
        nb = V * 4 * 4 * 4
        f = np.frombuffer(raw[:nb], ">f4").astype(np.float64).reshape(V, 4, 2, 2)
        z0 = f[:, :, 0, 0] + 1j * f[:, :, 0, 1]
        z1 = f[:, :, 1, 0] + 1j * f[:, :, 1, 1]
        W = np.zeros((V, 4, 2, 2), complex)
        W[:, :, 0, 0] = np.conj(z0)
        W[:, :, 1, 1] = z0
        W[:, :, 0, 1] = -z1
        W[:, :, 1, 0] = np.conj(z1)
        U[order] = W
        p = nb
        plaq, _, _, beta = np.frombuffer(raw[p:p + 16], ">f4").astype(float)
        p += 16
        kappas = []
        while True:
            k = float(np.frombuffer(raw[p:p + 4], ">f4")[0])
            p += 4
            kappas.append(k)
            if k >= 0:
                break
        nk = len(kappas)
        p += 4 * nk + 8 * nk
        sweep = int(np.frombuffer(raw[p:p + 4], ">i4")[0])
        p += 4 + 20 * nk
        lat = [int(v) for v in np.frombuffer(raw[p:p + 16], ">i4")]
        info = dict(plaquette=float(plaq), beta=float(beta), kappa=kappas[-1], sweep=sweep, dims=lat)
    else:
        nb = V * 4 * 6 * 8
        f = np.frombuffer(raw[:nb], ">f4").astype(np.float64).reshape(V, 4, 3, 2, 2)
        m = f[..., 0] - 1j * f[..., 1]
        U[order] = su3_reunitarise(m)
        p = nb
        lat = [int(v) for v in np.frombuffer(raw[p:p + 16], ">i4")]
        beta, c1, plaq = np.frombuffer(raw[p + 16:p + 40], ">f8")
        info = dict(plaquette=float(plaq), beta=float(beta), c1=float(c1), dims=lat)
    if info["dims"] != list(dims):
        raise ValueError("lattice size is different")
    return U, info
