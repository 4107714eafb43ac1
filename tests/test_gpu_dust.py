"""GPU parity of the dust-opacity generator (mcx_compute_parts) against the
oracle's ComputePart restatement.

The library computes every transcendental starting value on the host the
way the oracle does, and the kernels use only IEEE +, -, *, / in the
oracle's order without contraction: opacities, scattering matrices and the
mean density must be bit-identical, and both take the same paths (layered
spheres, homogeneous spheres for large grains, solid spheres and the
non-converged fallback)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _species(lam, nmat=1, **kw):
    nlam = len(lam)
    e1 = np.empty((nmat, nlam), np.float32)
    e2 = np.empty((nmat, nlam), np.float32)
    for l in range(nmat):
        # a smooth n and an absorption band in k, different per material
        e1[l] = 1.5 + 0.2 * l + 0.1 * np.log10(lam) / 3
        e2[l] = 0.01 + 0.3 * np.exp(-(np.log10(lam / (10.0 * (l + 1)))) ** 2 / 0.02) + 0.002 * lam ** 0.5
    frac = np.full(nmat, 1.0 / nmat, np.float32)
    rho = (3.0 + 0.5 * np.arange(nmat)).astype(np.float32)
    d = dict(e1=e1, e2=e2, frac=frac, rho=rho, amin=0.05, amax=20.0, apow=3.5, fmax=0.8, ns=4, nf=5)
    d.update(kw)
    return d


def _compare(gpu, oracle_lib, lam, parts, expect=None):
    res, st = gpu.compute_parts(lam, parts)
    tot = dict(ndhs=0, nhom=0, nerr=0)
    for p, r in zip(parts, res):
        o = oracle_lib.compute_part(lam, p["e1"], p["e2"], p["frac"], p["rho"], p["amin"], p["amax"], p["apow"],
                                    p["fmax"], p["ns"], p["nf"])
        for k in ("Kabs", "Ksca", "Kext", "F"):
            bad = np.argwhere(r[k] != o[k])
            assert len(bad) == 0, (k, bad[:5], r[k].ravel()[:3], o[k].ravel()[:3])
        assert r["rho"] == o["rho"]
        for k in tot:
            tot[k] += o[k]
    for k in tot:
        assert st[k] == tot[k], (k, st[k], tot[k])
    if expect:
        for k, v in expect.items():
            assert (st[k] > 0) == v, (k, st[k])
    print("stats", st)
    return res, st


def test_dhs_two_materials_bit_identical(gpu, oracle_lib):
    lam = np.logspace(-1, 3, 16)
    _compare(gpu, oracle_lib, lam, [_species(lam, nmat=2)], expect=dict(ndhs=True))


def test_several_species_one_call(gpu, oracle_lib):
    lam = np.logspace(-1, 3, 12)
    parts = [_species(lam, amin=0.05, amax=0.3, ns=3, nf=4), _species(lam, nmat=2, amin=0.3, amax=3.0),
             _species(lam, amin=3.0, amax=30.0, ns=2, nf=3)]
    _compare(gpu, oracle_lib, lam, parts)


def test_solid_spheres(gpu, oracle_lib):
    lam = np.logspace(-1, 3, 10)
    _compare(gpu, oracle_lib, lam, [_species(lam, fmax=0.0, ns=5)], expect=dict(ndhs=False, nhom=True))


def test_large_grains_use_homogeneous_spheres(gpu, oracle_lib):
    # 2 r/lam > 10^4/pi at the short wavelengths: MeerhoffMie, with the size
    # parameter capped at 2 pi 5000 below lam = r/5000
    lam = np.logspace(-1, 1, 6)
    _compare(gpu, oracle_lib, lam, [_species(lam, amin=600.0, amax=900.0, ns=2, nf=3)],
             expect=dict(ndhs=True, nhom=True))


def test_non_converged_layers_fall_back(gpu, oracle_lib):
    # |m| close to 1: the layered series needs more terms than x|m| and
    # the reference falls back to a homogeneous sphere from that fraction on
    lam = np.logspace(-1, 0.5, 6)
    p = _species(lam, amin=1.0, amax=40.0, ns=3, nf=4)
    p["e1"][:] = 1.02
    p["e2"][:] = 0.001
    _compare(gpu, oracle_lib, lam, [p], expect=dict(nerr=True, nhom=True))


def test_normalisation_and_cross_sections(gpu):
    """sum F11 sin = sum sin to single precision, Kext = Kabs + Ksca at the
    REAL level, positive absorption."""
    lam = np.logspace(-1, 3, 20)
    res, _ = gpu.compute_parts(lam, [_species(lam, nmat=2)])
    r = res[0]
    th = (np.arange(180) + 0.5) * np.pi / 180
    norm = (r["F"][:, 0, :] * np.sin(th)).sum(axis=1) / np.sin(th).sum()
    assert np.all(np.abs(norm - 1) < 2e-6), norm
    assert np.all(r["Kabs"] > 0) and np.all(r["Ksca"] > 0)
    assert np.allclose(r["Kext"], r["Kabs"] + r["Ksca"], rtol=2e-6)
    # a sum of single-particle Mueller matrices keeps F11 >= |F12|, |F33|,
    # |F34| (the first bin carries the forward-peak correction of F11 only)
    F = r["F"][:, :, 1:]
    for q in (1, 3, 4):
        assert np.all(np.abs(F[:, q]) <= F[:, 0] * (1 + 1e-5)), q


def test_production_dust_through_the_keywords(gpu, workdir, oracle_lib):
    """The production model's computepart particle (10 species, 10 sizes x
    10 DHS fractions each) built from its keyword file; sampled rows of three
    species against the oracle, and the MRN mass fractions of the split."""
    from mcmax_amd import configs, dust, keywords
    wd = workdir + "/gpu_prod_dust"
    m = configs.load_config("production", wd, tr=gpu)
    assert m.ng == 10
    st = m.dust_stats
    assert st["ndhs"] > 0 and st["nhom"] > 0 and st["nerr"] == 0
    print("production dust", st)
    kw = keywords.parse_file(wd + "/production.in")
    specs, w = dust.species_specs(dust.particle_keywords(kw)[1], m.lam, wd)
    assert np.allclose(m.w[1, 0], w / w.sum(), rtol=1e-12)
    sel = [0, 57, 120, 199]
    for s in (0, 6, 9):
        p = specs[s]
        o = oracle_lib.compute_part(m.lam[sel], p["e1"][:, sel], p["e2"][:, sel], p["frac"], p["rho"], p["amin"],
                                    p["amax"], p["apow"], p["fmax"], p["ns"], p["nf"])
        assert np.array_equal(m.Kabs[s][sel], o["Kabs"]) and np.array_equal(m.Ksca[s][sel], o["Ksca"])
        assert np.array_equal(m.F[s][sel], o["F"])
