"""Output products and observation files: the structure file round trip,
the spectrum writers' columns and names, and arrinterpol's behaviour."""
import math
import os

import numpy as np
import pytest

from mcmax_amd import iterate, observe, output


def test_angle_tag_matches_fortran_edit_descriptors():
    # i1 of int(angle/10) then f3.1 of the remainder (Telescope.f:145-147)
    assert output.angle_tag(45.0) == "45.0"
    assert output.angle_tag(5.5) == "05.5"
    assert output.angle_tag(89.04) == "89.0"
    assert output.angle_tag(0.0) == "00.0"
    # a remainder that rounds to 10.0 overflows the field
    assert output.angle_tag(39.97) == "3***"
    assert output.spectrum_filename(60.0, "_a") == "spectrum60.0_a.dat"
    assert output.spectrum_filename(60.0, "_a", noangle=True) == "spectrum_a.dat"


def test_arrinterpol_log_against_oracle_and_its_scan(oracle_lib):
    xs = np.array([1.0, 2.0, 4.0, 8.0, 16.0])
    ys = np.array([3.0, 5.0, 7.0, 11.0, 13.0])
    # outside the table: the abscissa at the end is returned (as the reference does)
    assert output.arrinterpol_log(0.5, xs, ys) == 1.0
    assert output.arrinterpol_log(20.0, xs, ys) == 16.0
    # inside: the scan keeps the LAST index with x <= xarr(i), so the power
    # law through the last two points is evaluated for every interior x
    pw = math.log10(13.0 / 11.0) / math.log10(2.0)
    assert output.arrinterpol_log(3.0, xs, ys) == pytest.approx(11.0 * (3.0 / 8.0) ** pw, rel=1e-14)
    lib = oracle_lib.load()
    rng = np.random.default_rng(3)
    xa = np.sort(rng.uniform(0.1, 100.0, 40))
    ya = rng.uniform(0.5, 5.0, 40)
    for x in rng.uniform(0.05, 120.0, 200):
        a = output.arrinterpol_log(float(x), xa, ya)
        b = lib.orc_arrinterpol_log(float(x), xa.ctypes.data_as(oracle_lib._dp), ya.ctypes.data_as(oracle_lib._dp),
                                    len(xa))
        assert a == pytest.approx(b, rel=1e-13, abs=0)


def test_denstemp_round_trip(tiny_model, tmp_path):
    m = tiny_model
    st = iterate.density_state(m)
    rng = np.random.default_rng(1)
    T = rng.uniform(10.0, 1500.0, (m.nr, m.nt))
    st["gasdens"] = st["dens"] * 0.7
    st["dens0"] = st["dens"] * 1.3
    p = tmp_path / "denstemp.dat"
    output.write_denstemp(str(p), m, T, st)
    d = output.read_denstemp(str(p))
    assert d["T"].shape == (m.nr - 1, m.nt)
    assert np.array_equal(d["R_av"], m.R_av[1:])
    assert np.array_equal(d["theta_av"], m.theta_av)
    assert np.array_equal(d["T"], T[1:])
    assert np.array_equal(d["dens"], st["dens"][1:])
    assert np.array_equal(d["dens0"], st["dens0"][1:])
    assert np.allclose(d["gasdens"], st["gasdens"][1:], rtol=1e-15, atol=0)
    assert np.array_equal(d["w"], m.w[1:])
    lines = p.read_text().splitlines()
    assert lines[0] == "# Format number" and int(lines[1]) == 5
    assert [int(x) for x in lines[3].split()] == [m.nr - 1, m.nt, m.ng, 1]


def test_denstemp_rejects_other_formats(tmp_path):
    p = tmp_path / "d.dat"
    p.write_text("# Format number\n     4\n")
    with pytest.raises(output.OutputError):
        output.read_denstemp(str(p))


def test_mcspec_files(tiny_model, tmp_path):
    m = tiny_model
    nang = 30
    rng = np.random.default_rng(2)
    ispec = rng.integers(0, 50, (nang + 1, m.nlam))
    tal = dict(spec=rng.uniform(0, 1e5, (nang + 1, m.nlam)), scatspec=rng.uniform(0, 1e4, (nang + 1, m.nlam)),
               ispec=ispec, iscatspec=ispec // 2)
    paths = output.write_mcspec(str(tmp_path), m, tal)
    assert len(paths) == nang + 1
    # MCRadiation.f:740: angle(j) = acos((j - 0.5)/nangle) in degrees
    names = {os.path.basename(p) for p in paths}
    assert "MCSpec89.0.dat" in names and "MCSpec.dat" in names
    j = 7
    ang = 180.0 * math.acos((j - 0.5) / nang) / math.pi
    rows = np.loadtxt(tmp_path / f"MCSpec{output.angle_tag(ang)}.dat")
    d2 = m.distance_cm ** 2
    assert rows.shape == (m.nlam, 6)
    assert np.allclose(rows[:, 0], m.lam, rtol=1e-15)
    assert np.allclose(rows[:, 1], tal["spec"][j] / d2, rtol=1e-15)
    ok = ispec[j] > 0
    dspec = tal["spec"][j] / np.sqrt(ispec[j].astype(np.float32)).astype(np.float64)
    assert np.allclose(rows[ok, 2], dspec[ok] / d2, rtol=1e-15)
    assert np.array_equal(rows[:, 4], tal["iscatspec"][j])
    assert np.allclose(rows[:, 5], 1e23 * m.Fstar / d2, rtol=1e-15)
    tot = np.loadtxt(tmp_path / "MCSpec.dat")
    assert np.allclose(tot[:, 1], tal["spec"][0] / d2, rtol=1e-15)


def test_spectrum_file_columns(tiny_model, tmp_path):
    m = tiny_model
    lam = np.array([0.5, 2.0, 10.0])
    fl = np.array([1e30, 2e30, 3e30])
    p = tmp_path / "spectrum45.0.dat"
    output.write_spectrum(str(p), m, lam, fl, fl / 10, fl / 100)
    rows = np.loadtxt(p)
    d2 = m.distance_cm ** 2
    assert rows.shape == (3, 5 if m.scat_how == 2 else 4)
    assert np.allclose(rows[:, 1], 1e23 * fl / d2, rtol=1e-15)
    assert np.allclose(rows[:, 2], 1e23 * fl / 10 / d2, rtol=1e-15)
    fstar = [output.arrinterpol_log(x, m.lam, m.Fstar) for x in lam]
    assert np.allclose(rows[:, 3], 1e23 * np.array(fstar) / d2, rtol=1e-15)


def test_observation_file(tiny_model, tmp_path):
    m = tiny_model
    p = tmp_path / "obs.in"
    p.write_text("* comment\nnphot=200\ntype='SPECTRUM'\ninc=30\nnlam=7\nlam1=0.5\nlam2=500\n"
                 "type=SPECTRUM\nflag='_b'\n")
    obs = observe.parse_observation_file(str(p), m)
    assert len(obs) == 2
    a, b = obs
    assert a["inc"] == 30.0 and a["nphot"] == 200 and b["nphot"] == 200
    assert b["inc"] == m.incangle and b["flag"] == "_b" and b["nphi"] == 45
    la = observe.observed_wavelengths(a, m.lam)
    assert len(la) == 7 and la[0] == pytest.approx(0.5) and la[-1] == pytest.approx(500.0, rel=1e-6)
    assert np.allclose(np.diff(np.log10(la)), math.log10(1000.0) / 6, rtol=1e-6)
    lb = observe.observed_wavelengths(b, m.lam)
    # the model grid with both ends (Telescope.f:40-71)
    assert lb[0] == m.lam[0] and lb[-1] == m.lam[-1]
    assert np.array_equal(lb[1:-1], m.lam[1:-1])
    p.write_text("type=VISIBILITY\n")
    with pytest.raises(observe.ObservationError):
        observe.parse_observation_file(str(p), m)


def _parse_fits_independently(raw):
    """A separate reading of the FITS rules: 80-column cards, '= ' in columns
    9-10, END, 2880-byte blocks, big-endian doubles."""
    import re
    assert len(raw) % 2880 == 0
    pos, out = 0, []
    while pos < len(raw):
        cards = {}
        while True:
            blk = raw[pos:pos + 2880]
            pos += 2880
            lines = [blk[i:i + 80].decode("ascii") for i in range(0, 2880, 80)]
            end = False
            for ln in lines:
                assert len(ln) == 80
                if ln.startswith("END     "):
                    end = True
                    break
                mo = re.match(r"^([A-Z0-9_-]{1,8})\s*= +(.*)$", ln)
                if mo:
                    cards[mo.group(1)] = mo.group(2).split(" / ")[0].strip()
            if end:
                break
        dims = [int(cards[f"NAXIS{k}"]) for k in range(1, int(cards["NAXIS"]) + 1)]
        n = int(np.prod(dims))
        data = np.frombuffer(raw[pos:pos + 8 * n], dtype=">f8").reshape(dims[::-1])
        pos += 8 * n
        pos += (-pos) % 2880
        out.append((cards, data))
    return out


def test_fits_structure_file(tiny_model, tmp_path):
    import gzip
    m = tiny_model
    st = iterate.density_state(m)
    rng = np.random.default_rng(5)
    T = rng.uniform(10.0, 900.0, (m.nr, m.nt))
    st["gasdens"] = st["dens"] * 0.5
    p = tmp_path / "denstemp.fits.gz"
    output.write_denstemp_fits(str(p), m, T, st)
    raw = gzip.open(p).read()
    hdus = _parse_fits_independently(raw)
    assert len(hdus) == 6
    h0, grid = hdus[0]
    assert h0["SIMPLE"] == "T" and h0["BITPIX"] == "-64" and h0["EXTEND"] == "T"
    assert int(h0["NR"]) == m.nr - 1 and int(h0["NTHETA"]) == m.nt and int(h0["NGRAINS"]) == m.ng
    assert [hdus[k][0]["XTENSION"].strip("' ") for k in range(1, 6)] == ["IMAGE"] * 5
    # Fortran a(i, j) with i (radius) fastest: numpy [j, i]
    from mcmax_amd.model import AU
    assert np.array_equal(grid[0, 3, :], m.R_av[1:] / AU)
    assert np.array_equal(grid[1, :, 2], m.theta_av)
    assert np.array_equal(hdus[2][1], T[1:].T)
    assert hdus[3][1].shape == (2, m.ng, m.nt, m.nr - 1)
    assert np.array_equal(hdus[4][1], (st["gasdens"] * 100.0)[1:].T)
    # and the library's own reader gives back the ASCII reader's dict
    d = output.read_denstemp(str(p))
    assert np.array_equal(d["T"], T[1:]) and np.array_equal(d["dens"], st["dens"][1:])
    assert np.allclose(d["gasdens"], st["gasdens"][1:], rtol=1e-15, atol=0)
    assert np.array_equal(d["w"], m.w[1:])
