"""CPU: MAOS .conf reading (mt_conf_*, maos_amd.conf).

The value parsers against the examples the reference documents for them
(sys/readstr.c:30-39 for string arrays as the code parses them, :177-192
for numeric arrays); the key store's include / override / append /
continuation rules; and a configuration written as MAOS keys
(tests/golden/conf) that gives the same setup, bit for bit, as the
benchmark configuration it describes.
"""
import os

import numpy as np
import pytest

from maos_amd import _abi, configs, tomo
from maos_amd.conf import Conf, fit_config, tomo_config

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "conf")


def conf_of(text):
    c = Conf()
    c.load(text)
    return c


@pytest.mark.parametrize("value,want,shape", [
    ("[1 2 3]", [1, 2, 3], (3, 1)),
    ("[1 2 3]+2", [3, 4, 5], (3, 1)),
    ("2*[1 2 3]", [2, 4, 6], (3, 1)),
    ("[1 2 3]*2", [2, 4, 6], (3, 1)),
    ("[1 2 3]/2", [0.5, 1, 1.5], (3, 1)),
    ("2/[1 2 3]", [2, 1, 2 / 3], (3, 1)),
    ("2/[1 2 3]*2+1", [5, 3, 7 / 3], (3, 1)),
    ("[1 2 3; 4 5 6]", [1, 2, 3, 4, 5, 6], (3, 2)),
    ("[1 2 3; 4 5 6]'", [1, 4, 2, 5, 3, 6], (2, 3)),
    ("[1/2 1/12 0.5*3]", [0.5, 1 / 12, 1.5], (3, 1)),
    ("[0 1 2]*1000", [0, 1000, 2000], (3, 1)),
    ("[inf]", [np.inf], (1, 1)),
    ("0.5", [0.5], (1, 1)),
    ("[]", [], (0, 0)),
])
def test_numeric_arrays(value, want, shape):
    c = conf_of("a = %s" % value)
    a = c.arr2("a")
    assert a.shape == shape
    assert np.allclose(a.ravel(order="F"), want, rtol=1e-15, atol=0)


@pytest.mark.parametrize("value,want", [
    ("[a b]", ["a", "b"]),
    ("[a,b]", ["a", "b"]),
    ("['a b']", ["a b"]),
    ("[,]", ["", ""]),
    ("[\"x.conf\" 'y z', w]", ["x.conf", "y z", "w"]),
])
def test_string_arrays(value, want):
    assert conf_of("s = %s" % value).strarr("s") == want


def test_bad_values_are_errors():
    c = conf_of("a = [1 2; 3]\nb = [1 2]x\nq = \"abc")
    for k in ("a", "b"):
        with pytest.raises(_abi.MtError):
            c.arr(k)
    with pytest.raises(_abi.MtError):
        c.str("q")
    with pytest.raises(_abi.MtError):
        c.arr("missing")


def test_store_rules(tmp_path):
    (tmp_path / "base.conf").write_text("x = 1\ny = [1 2]\nz = 'quoted value'  # comment\n")
    (tmp_path / "top.conf").write_text("include = \"base.conf\"\nx = 2\ny += [3]\nw = [1 \\\n 2 \\\n 3]\n")
    c = Conf.read(str(tmp_path / "top.conf"))
    assert c.num("x") == 2
    assert list(c.arr("y")) == [1, 2, 3]
    assert c.str("z") == "quoted value"
    assert list(c.arr("w")) == [1, 2, 3]
    # a lower priority does not override, an equal or higher one does
    c.load("x = 5", priority=0)
    c.load("x = 7", priority=1)
    c.load("x = 9", priority=0)
    assert c.num("x") == 7
    # keys never read
    c2 = Conf.read(str(tmp_path / "top.conf"))
    c2.num("x")
    assert set(c2.unused()) == {"y", "z", "w"}
    # __reset__ drops everything read so far
    c2.load("__reset__\nq = 1")
    assert not c2.has("x") and c2.has("q")


def test_bin_file_value(tmp_path):
    from maos_amd import binio
    a = np.arange(6.0).reshape(3, 2, order="F")
    binio.write(str(tmp_path / "v.bin"), a)
    c = Conf()
    c.add_path(str(tmp_path))
    c.load("m = v.bin")
    got = c.arr2("m")
    assert np.array_equal(got, a)


def test_large30_from_conf_equals_config():
    c = Conf.read(os.path.join(GOLD, "large30.conf"))
    cfg = tomo_config(c, name="large30")
    ref = configs.large30()
    for k in ("aper_d", "aper_din", "ht", "wt", "os", "atmr_dx", "r0z", "L0", "tomo_pos", "tomo_guard", "tomo_cone",
              "tomo_precond", "recon_split", "tomo_splitlrt", "maxit"):
        assert cfg[k] == ref[k], k
    assert cfg["powfs"] == ref["powfs"] and cfg["wfs"] == ref["wfs"] and cfg["dirs"] == ref["dirs"]
    assert c.unused() == [] or set(c.unused()) <= {"atm.frozenflow"}
    a, b = tomo.Recon(cfg), tomo.Recon(ref)
    assert a.nx == b.nx and a.ng == b.ng
    fa, fb = a.fdpcg(), b.fdpcg()
    assert np.array_equal(fa["Minv"], fb["Minv"])
    assert np.array_equal(a.l2c(), b.l2c())


def test_fit_from_conf_equals_config():
    c = Conf.read(os.path.join(GOLD, "large30.conf"), os.path.join(GOLD, "fit_dual.conf"))
    fc = fit_config(c)
    ref = configs.fit_for(configs.large30())
    assert fc["dm"] == ref["dm"]
    assert np.allclose([f[:2] for f in fc["fit"]], [f[:2] for f in ref["fit"]], rtol=0, atol=1e-12)
    assert np.allclose([f[2] for f in fc["fit"]], [f[2] for f in ref["fit"]], rtol=1e-15)
    base = tomo_config(c, name="large30")
    assert len(base["dirs"]) == 1 + 9
    r = tomo.Recon(configs.with_fit_dirs(configs.large30(), ref))
    r2 = tomo.Recon(base)
    assert r.nx == r2.nx
    f1, f2 = tomo.FitSetup(r2, fc), tomo.FitSetup(r, ref)
    assert f1.nact == f2.nact and f1.rec.maxit == 100
    assert np.array_equal(f1.W1(), f2.W1())


REFCONF = "/root/reference/config/maos"


@pytest.mark.skipif(not os.path.isdir(REFCONF), reason="reference configuration files not present")
def test_reference_configuration_files_parse():
    """Every .conf file of the reference's configuration directory reads,
    and every value in it parses as a number array or a string array."""
    n = 0
    for fn in sorted(os.listdir(REFCONF)):
        if not fn.endswith(".conf"):
            continue
        c = Conf()
        c.add_path(REFCONF)
        c.load(fn)
        n += 1
    assert n > 20
    c = Conf.read("default.conf", paths=[REFCONF])
    assert c.arr("aper.d")[0] > 0
    assert abs(np.sum(Conf.read("fit_cir30.conf", paths=[REFCONF]).arr("fit.wt")) - 1) < 1e-12
