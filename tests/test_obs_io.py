"""Bin files of the equal-time measurements (the reference's text format) and their
jackknife analysis: formatting, round trips and closed-form checks; no device needed."""
import os

import numpy as np

from alf_amd import obs_io
from helpers import hubbard


def test_fortran_e_format():
    assert obs_io.fortran_e(1.0) == " 0.10000000000000000E+001"
    assert obs_io.fortran_e(-2.0 ** -22) == "-0.23841857910156250E-006"
    assert obs_io.fortran_e(2.5e-7) == " 0.24999999999999999E-006"      # 2.5e-7 is not a binary fraction
    assert obs_io.fortran_e(0.0) == " 0.00000000000000000E+000"
    assert obs_io.fortran_e(123456.789, 26) == "  0.12345678900000000E+006"
    assert len(obs_io.fortran_e(-1.2345678901234567e-300)) == 25
    for x in (3.141592653589793, -1e-12, 7.5e10):
        assert float(obs_io.fortran_e(x).replace("E", "e")) == x


def fake_bin(ham, rng, n=10):
    nc, no = ham.latt.N, ham.unit.Norb
    return {"scal": rng.standard_normal(4) * n + 1j * rng.standard_normal(4), "sign_scal": float(n), "n_scal": n,
            "latt": (rng.standard_normal((3, nc, no, no)) + 1j * rng.standard_normal((3, nc, no, no))) * n,
            "den0": rng.standard_normal(no) * n * nc + 0j, "sign_eq": float(n), "n_eq": n}


def test_scalar_bins_round_trip_and_jackknife(tmp_path):
    ham = hubbard(L1=4, L2=4)
    w = obs_io.BinWriter(str(tmp_path), ham)
    rng = np.random.default_rng(0)
    bins = [fake_bin(ham, rng) for _ in range(6)]
    for b in bins:
        w.write(b)
    assert open(tmp_path / "Kin_scal_info").read().startswith("====== Analysis Mode ======")
    for q, name in enumerate(obs_io.SCALARS):
        vals, signs = obs_io.read_scal(str(tmp_path / f"{name}_scal"))
        want = np.array([b["scal"][q] / b["n_scal"] for b in bins])
        assert vals.shape == (6, 1) and np.allclose(signs, 1.0)
        assert np.allclose(vals[:, 0], want, rtol=1e-15, atol=0)
        (best, err), = obs_io.jackknife_scal(str(tmp_path / f"{name}_scal"))
        x = want.real
        assert abs(best.real - x.mean()) < 1e-12
        # unit sign: jackknife values (sum - x_i) / (N - 1) deviate by (mean - x_i) / (N - 1)
        assert abs(err.real - np.sqrt(np.sum((x - x.mean()) ** 2)) / (len(x) - 1)) < 1e-12


def test_jackknife_of_a_ratio_with_sign():
    rng = np.random.default_rng(3)
    o = rng.standard_normal(8) + 2.0
    s = rng.uniform(0.5, 1.0, 8)
    best, err = obs_io.jackknife(lambda X: X[0] / X[1], np.vstack([o, s]).astype(complex))
    j = [(o.sum() - o[i]) / (s.sum() - s[i]) for i in range(8)]
    assert abs(best.real - np.mean(j)) < 1e-13 and abs(err.real - np.sqrt(np.sum((j - np.mean(j)) ** 2))) < 1e-13
    b2, _ = obs_io.jackknife(lambda X: X[0] / X[1], np.vstack([o, s]).astype(complex), n_skip=2, n_rebin=3)
    r = [(o[2 + 3 * i:5 + 3 * i].mean(), s[2 + 3 * i:5 + 3 * i].mean()) for i in range(2)]
    jj = [(r[1 - i][0]) / (r[1 - i][1]) for i in range(2)]
    assert abs(b2.real - np.mean(jj)) < 1e-13


def test_correlation_bins_layout_and_transform(tmp_path):
    """A correlation that lives only at distance class 0 (R = 0) is flat in k: X(k) = x0 / N."""
    ham = hubbard(L1=3, L2=3, lattice="Honeycomb")
    from alf_amd.host import lattices as LT
    w = obs_io.BinWriter(str(tmp_path), ham)
    nc, no = ham.latt.N, ham.unit.Norb
    imj = LT.distance_classes(ham.latt)
    d0 = int(imj[0, 0])
    latt = np.zeros((3, nc, no, no), dtype=complex)
    latt[:, d0] = np.arange(1, 1 + 3 * no * no).reshape(3, no, no)
    b = {"scal": np.ones(4, complex), "sign_scal": 1.0, "n_scal": 1, "latt": latt,
         "den0": np.array([nc * 0.5, nc * 0.25]), "sign_eq": 1.0, "n_eq": 1}
    w.write(b)
    w.write(b)
    for q, name in enumerate(obs_io.CORRELATIONS):
        lines = open(tmp_path / f"{name}_eq").read().splitlines()
        per = 1 + no + nc * (1 + no * no)
        assert len(lines) == 2 * per
        head = lines[0].split()
        assert int(head[1]) == no and int(head[2]) == nc and float(head[0].replace("E", "e")) == 1.0
        bg = [complex(*map(float, l.strip("()").replace("E", "e").split(","))) for l in lines[1:1 + no]]
        assert np.allclose(bg, [0.5, 0.25] if name == "Den" else [0, 0])
        for ik in range(nc):
            base = 1 + no + ik * (1 + no * no)
            vals = [complex(*map(float, l.strip("()").replace("E", "e").split(",")))
                    for l in lines[base + 1:base + 1 + no * no]]
            assert np.allclose(vals, (np.arange(1, 1 + 3 * no * no).reshape(3, no * no)[q]) / nc)
    info = open(tmp_path / "Green_eq_info").read()
    assert "Observable: Green_eq" in info and "Number of orbitals:          2" in info


def test_tau_bins_layout_and_transform(tmp_path):
    """Green_tau / SpinZ_tau / Den_tau: header with Ntau and dtau, Norb background lines
    (Den0 / (N Ns Ntau)), then per momentum its coordinates and Ntau x Norb x Norb values; a
    correlation that lives only at R = 0 is flat in k for every tau."""
    ham = hubbard(L1=3, L2=3, lattice="Honeycomb", beta=0.3, dtau=0.1)
    from alf_amd.host import lattices as LT
    w = obs_io.BinWriter(str(tmp_path), ham)
    nc, no, nt = ham.latt.N, ham.unit.Norb, ham.Ltrot + 1
    d0 = int(LT.distance_classes(ham.latt)[0, 0])
    latt = np.zeros((3, nt, nc, no, no), dtype=complex)
    for t in range(nt):
        latt[:, t, d0] = (t + 1) * np.arange(1, 1 + 3 * no * no).reshape(3, no, no)
    sums = {"latt": 2 * latt, "den0": np.array([2 * nc * nt * 0.5, 2 * nc * nt * 0.25]), "n": 2, "sign": 2.0}
    w.write_tau(sums)
    for q, name in enumerate(obs_io.CORRELATIONS):
        lines = open(tmp_path / f"{name}_tau").read().splitlines()
        assert len(lines) == 1 + no + nc * (1 + nt * no * no)
        head = lines[0].split()
        assert [int(v) for v in head[1:4]] == [no, nc, nt]
        assert abs(float(head[4].replace("E", "e")) - 0.1) < 1e-15
        assert float(head[0].replace("E", "e")) == 1.0
        bg = [complex(*map(float, l.strip("()").replace("E", "e").split(","))) for l in lines[1:1 + no]]
        assert np.allclose(bg, [0.5, 0.25] if name == "Den" else [0, 0])
        for ik in range(nc):
            base = 1 + no + ik * (1 + nt * no * no)
            vals = np.array([complex(*map(float, l.strip("()").replace("E", "e").split(",")))
                             for l in lines[base + 1:base + 1 + nt * no * no]]).reshape(nt, no * no)
            for t in range(nt):
                assert np.allclose(vals[t], (t + 1) * np.arange(1, 1 + 3 * no * no).reshape(3, no * no)[q] / nc)
        info = open(tmp_path / f"{name}_tau_info").read()
        assert f"Observable: {name}_tau" in info and f"{'Ntau':>20s}: {nt:10d}" in info
        assert ("Channel: P\n" if name == "Green" else "Channel: PH\n") in info
