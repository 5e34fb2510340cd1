"""Cross-section files beyond read_cross_section (mcc.cu:17-41): the m*10^{e}
notation of csf/Ar_conv.txt, LXCat blocks loaded into a cross-section set,
read-back of the processes (host code only, no GPU)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_pow10_notation_reads_the_same_doubles(pic):
    # csf/Ar_conv.txt holds the rows of csf/Ar.txt written as 6.2984*10^{-20}
    e, s = pic.read_table(os.path.join(GOLDEN, "csf_Ar_conv_head.txt"))
    e_ref, s_ref = pic.read_cross_section(os.path.join(GOLDEN, "csf_Ar.txt"))
    assert len(e) == 20
    assert np.array_equal(e, e_ref[:20]) and np.array_equal(s, s_ref[:20])
    # on e-notation files both readers agree
    e2, s2 = pic.read_table(os.path.join(GOLDEN, "csf_Ar.txt"))
    assert np.array_equal(e2, e_ref) and np.array_equal(s2, s_ref)


def test_pow10_forms(pic, tmp_path):
    p = tmp_path / "t.txt"
    p.write_text("1.5*10^3\t2*10^{-21}\n2.5*10^{+3} 3.25e-21\n# c\n3*10^{x} 1\n4*10^{-2 1\n")
    e, s = pic.read_table(str(p))
    assert list(e) == [1500.0, 2500.0] and list(s) == [2e-21, 3.25e-21]


def _block(kind, target, param, rows, pow10=False):
    def num(v):
        if not pow10:
            return "%.6e" % v
        m, x = ("%.6e" % v).split("e")
        return "%s*10^{%s}" % (m, x)
    head = [kind, target] + ([] if param is None else [num(param) if not isinstance(param, str) else param])
    text = head + ["SPECIES: e / Ar", "PROCESS: E + Ar -> %s" % target, "COMMENT: synthetic test block",
                   "COLUMNS: Energy (eV) | Cross section (m2)", "-----------------------------"]
    text += ["%s\t%s" % (num(a), num(b)) for a, b in rows]
    text += ["-----------------------------", ""]
    return text


def _lxcat_file(tmp_path, crlf=False, extra=()):
    el_e, el_s = np.loadtxt(os.path.join(GOLDEN, "csf_Ar.txt"), unpack=True)
    io_e, io_s = np.loadtxt(os.path.join(GOLDEN, "csf_Ar+.txt"), unpack=True)
    exc = [(11.5, 0.0), (12.0, 1e-21), (20.0, 8e-21), (100.0, 5e-21)]
    lines = ["LXCat, www.lxcat.net", "Generated on 1 Jan 2000. All rights reserved.", "",
             "DATABASE:         synthetic", "*" * 40, ""]
    lines += _block("ELASTIC", "Ar", 1.36e-5, list(zip(el_e, el_s)))
    lines += _block("EXCITATION", "Ar -> Ar*(11.5eV)", "1.150000e+1  1.0", exc, pow10=True)
    lines += _block("ATTACHMENT", "Ar -> Ar^-", None, [(0.0, 0.0), (10.0, 1e-22)])
    lines += _block("IONIZATION", "Ar -> Ar^+", float(io_e[0]), list(zip(io_e, io_s)))
    for blk in extra:
        lines += blk
    lines += ["xxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxxx"]
    p = tmp_path / "ar.lxcat"
    p.write_bytes(("\r\n" if crlf else "\n").join(lines).encode())
    return str(p), (el_e, el_s), exc, (io_e, io_s)


@pytest.mark.parametrize("crlf", [False, True])
def test_lxcat_blocks_become_processes(pic, tmp_path, crlf):
    path, (el_e, el_s), exc, (io_e, io_s) = _lxcat_file(tmp_path, crlf)
    xs = pic.CrossSections(1e21, 0.03)
    xs.add_analytic(1, pic.CHARGE_EXCHANGE, 0.0, 5e-19)
    added, skipped = xs.load_lxcat(path, 0, product_species=1)
    assert (added, skipped) == (3, 1)
    procs = xs.processes()
    assert [(p["species"], p["type"]) for p in procs] == [
        (1, pic.CHARGE_EXCHANGE), (0, pic.ELASTIC), (0, pic.EXCITATION), (0, pic.IONIZATION)]
    assert procs[0]["peak_m2"] == 5e-19 and len(procs[0]["energies"]) == 0
    el, ex, io = procs[1:]
    assert np.array_equal(el["energies"], el_e) and np.array_equal(el["sigmas"], el_s)
    assert ex["threshold_eV"] == 11.5 and ex["energies"].tolist() == [a for a, _ in exc]
    assert ex["sigmas"].tolist() == [b for _, b in exc]
    assert io["threshold_eV"] == 15.76 and io["product_species"] == 1
    assert np.array_equal(io["energies"], io_e) and np.array_equal(io["sigmas"], io_s)


def test_lxcat_rejects_whole_file(pic, tmp_path):
    xs = pic.CrossSections(1e21, 0.03)
    before = len(xs.processes())
    # a fourth and fifth process for species 0
    two_more = [_block("EFFECTIVE", "Ar", 1.36e-5, [(0.0, 1e-20), (1.0, 2e-20)]),
                _block("BACKSCAT", "Ar", 1.36e-5, [(0.0, 1e-20), (1.0, 2e-20)])]
    path, *_ = _lxcat_file(tmp_path, extra=two_more)
    with pytest.raises(pic.PicError, match="more than 4"):
        xs.load_lxcat(path, 0, product_species=1)
    # ionization without a product species
    path, *_ = _lxcat_file(tmp_path)
    with pytest.raises(pic.PicError, match="product species"):
        xs.load_lxcat(path, 0)
    # energies not increasing
    bad = tmp_path / "bad.lxcat"
    bad.write_text("\n".join(_block("ELASTIC", "Ar", 1e-5, [(1.0, 1e-20), (1.0, 2e-20)])))
    with pytest.raises(pic.PicError, match="increase"):
        xs.load_lxcat(str(bad), 0)
    # unterminated table
    cut = tmp_path / "cut.lxcat"
    cut.write_text("\n".join(_block("ELASTIC", "Ar", 1e-5, [(0.0, 1e-20), (1.0, 2e-20)])[:-2]))
    with pytest.raises(pic.PicError, match="not closed"):
        xs.load_lxcat(str(cut), 0)
    with pytest.raises(pic.PicError):
        xs.load_lxcat(str(tmp_path / "missing.lxcat"), 0)
    assert len(xs.processes()) == before


def test_lxcat_process_feeds_the_majorant_like_add_process(pic, tmp_path):
    """A loaded block and the same table given to add_process are the same process."""
    path, (el_e, el_s), exc, (io_e, io_s) = _lxcat_file(tmp_path)
    a = pic.CrossSections(1e21, 0.03)
    a.load_lxcat(path, 0, product_species=1)
    b = pic.CrossSections(1e21, 0.03)
    b.add_process(0, pic.ELASTIC, 0.0, el_e, el_s)
    b.add_process(0, pic.EXCITATION, 11.5, *zip(*exc))
    b.add_process(0, pic.IONIZATION, 15.76, io_e, io_s, product_species=1)
    for p, q in zip(a.processes(), b.processes()):
        for k in ("species", "type", "product_species", "threshold_eV"):
            assert p[k] == q[k], k
        assert np.array_equal(p["energies"], q["energies"]) and np.array_equal(p["sigmas"], q["sigmas"])


ME, M_AR, QE = 9.1093837015e-31, 39.948 * 1.66053906660e-27, 1.602176634e-19


def _real_ar(pic, tmp_path, points=2048, emax=200.0):
    path, *_ = _lxcat_file(tmp_path)
    xs = pic.CrossSections(1e21, 0.03, points=points, emax_eV=emax)
    xs.load_lxcat(path, 0, product_species=1)
    return xs


def test_table_resamples_irregular_tables(pic, tmp_path):
    """The uniform table of n_gas*sigma is the clamped linear interpolation
    of the file's irregular rows (numpy's interp as the check)."""
    xs = _real_ar(pic, tmp_path)
    procs = xs.processes()
    info = xs.build(0, [ME, M_AR], 1e-11)
    K = info["K"]
    assert (info["n_proc"], K, info["type"], info["threshold"]) == (3, 2048, [0, 1, 2], [0.0, 11.5, 15.76])
    E = np.arange(K) * (200.0 / (K - 1))
    tab = info["table"].reshape(3, K)
    for q in range(3):
        want = 1e21 * np.interp(E, procs[q]["energies"], procs[q]["sigmas"])
        np.testing.assert_allclose(tab[q], want, rtol=1e-12, atol=0.0)
    assert info["product_species"] == 1
    assert info["efac"] == 0.5 * ME / QE


def _nu(info, m, E):
    K, de = info["K"], 1.0 / info["inv_de"]
    tab = info["table"].reshape(info["n_proc"], K)
    x = E / de
    j = np.floor(x).astype(np.int64)
    t = x - j
    hi = j >= K - 1
    j[hi], t[hi] = K - 2, 1.0
    v = np.sqrt(2.0 * E * QE / m)
    nu = np.zeros_like(E)
    for q in range(info["n_proc"]):
        a, b = tab[q, j], tab[q, j + 1]
        nu += np.where(E >= info["threshold"][q], (a + t * (b - a)) * v, 0.0)
    return nu


@pytest.mark.parametrize("points,emax", [(2048, 200.0), (257, 60.0)])
def test_band_majorants_bound_the_real_argon_rate(pic, tmp_path, points, emax):
    """Null-collision correctness: in every energy band the majorant is at
    least the collision frequency of the interpolated tables at every energy
    (1e6 random energies plus the table nodes), and within 1% of its maximum."""
    xs = _real_ar(pic, tmp_path, points, emax)
    info = xs.build(0, [ME, M_AR], 1e-11)
    rng = np.random.default_rng(0)
    de = 1.0 / info["inv_de"]
    for b in range(16):
        lo = 0.0 if b == 0 else emax * 2.0 ** (b - 15)
        hi = emax if b == 15 else emax * 2.0 ** (b - 14)
        E = np.concatenate([rng.uniform(lo, hi, 62500), np.arange(np.ceil(lo / de), np.floor(hi / de) + 1) * de,
                            [lo, np.nextafter(hi, 0)]])
        E = E[(E >= lo) & (E <= hi)]
        nu = _nu(info, ME, E)
        maj = info["band_nu_max"][b]
        assert nu.max() <= maj, (b, nu.max(), maj)
        assert maj <= 1.01 * nu.max(), (b, nu.max(), maj)
        assert info["band_p_null"][b] == -np.expm1(-maj * 1e-11)
    assert info["nu_max"] == max(info["band_nu_max"])
