"""Host tools around the .deb driver, no GPU: the xyz -> deb converter (Matlab/xyz2deb.m)
against the reference's own Data/UH3_7nm.xyz and Data/UH3_7nm_reduced.xyz -> Data/UH3_7nm.deb
(byte for byte), its grouping rules on small hand-made inputs, and the .debI averaging
(Matlab/plotdebi.m) against scipy's not-a-knot cubic spline and numpy's trapezoid rule."""
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import cudebye_amd as cd

CLI = os.path.join(os.path.dirname(cd.__file__), "bin", "debye")


@pytest.mark.parametrize("name", ["UH3_7nm.xyz", "UH3_7nm_reduced.xyz"])
def test_xyz_to_deb_reproduces_reference_deb(golden_dir, tmp_path, name):
    """Six columns (element x y z occupancy Beq) and four (element x y z): the same .deb."""
    src = tmp_path / name
    shutil.copy(golden_dir / name, src)
    out = cd.xyz_to_deb(src)
    assert out == str(src)[:-3] + "deb"
    assert open(out, "rb").read() == open(golden_dir / "UH3_7nm.deb", "rb").read()


def test_cli_xyz2deb(golden_dir, tmp_path):
    src = tmp_path / "UH3_7nm_reduced.xyz"
    shutil.copy(golden_dir / "UH3_7nm_reduced.xyz", src)
    subprocess.run([CLI, "--xyz2deb", str(src)], check=True)
    assert open(tmp_path / "UH3_7nm_reduced.deb", "rb").read() == open(golden_dir / "UH3_7nm.deb", "rb").read()


def _deb_groups(path):
    lines = open(path).read().split("\n")
    S = int(lines[3])
    groups = [lines[4 + i].split() for i in range(S)]
    pos = np.array([[float(v) for v in l.split("\t")] for l in lines[4 + S:] if l])
    return lines[:4], groups, pos


def test_xyz_grouping_order_and_centring(tmp_path):
    """Subgroups: runs of equal (element, occupancy, Beq) after a stable sort by element name
    (byte order), occupancy, Beq; atom order kept inside a subgroup; each axis centred on the
    middle of its range; Rmax = 2 ceil(1.025 max r); header from the given values."""
    rows = [("Zn", 1.0, 2.0, 3.0, 1.0, 0.5), ("O", 0.0, 0.0, 0.0, 0.5, 0.0), ("Zn", -1.0, 0.5, 3.0, 1.0, 0.5),
            ("O", 4.0, 1.0, 1.0, 1.0, 0.0), ("O", 2.0, -2.0, 0.5, 0.5, 0.0), ("Zn", 0.5, 0.5, 0.5, 1.0, 0.25),
            ("Al", 3.0, 3.0, 3.0, 1.0, 0.0)]
    src = tmp_path / "m.xyz"
    src.write_text("".join(f"{e}  {x:.6f}  {y:.6f}  {z:.6f}\t{o}\t{b}\r\n" for e, x, y, z, o, b in rows))
    out = cd.xyz_to_deb(src, deb=tmp_path / "m.deb", rstep=0.001, wavelength=0.7093, hist=2)
    head, groups, pos = _deb_groups(out)
    assert head[0] == "0 16 2048" and head[1] == "0.709300 10.00 140.00 0.02"
    # expected order: Al; O occ 0.5 (two atoms, file order); O occ 1; Zn Beq 0.25; Zn Beq 0.5 (two)
    want = [("Al", 1.0, 0.0, [6]), ("O", 0.5, 0.0, [1, 4]), ("O", 1.0, 0.0, [3]), ("Zn", 1.0, 0.25, [5]),
            ("Zn", 1.0, 0.5, [0, 2])]
    assert [g[1] for g in groups] == [w[0] for w in want]
    assert [int(g[0]) for g in groups] == [len(w[3]) for w in want]
    assert [float(g[2]) for g in groups] == [w[1] for w in want]
    assert [float(g[3]) for g in groups] == [w[2] for w in want]
    xyz = np.array([r[1:4] for r in rows])
    c = xyz - xyz.min(0) - (xyz.max(0) - xyz.min(0)) / 2
    order = [i for w in want for i in w[3]]
    assert np.allclose(pos, c[order], atol=5e-7)
    rmax = np.sqrt((c ** 2).sum(1)).max()
    assert head[2] == f"{2 * int(np.ceil(rmax * 1.025))} 0.0010 2"
    # the coefficients are the table's: O row of data/atomic_scattering_factors.txt
    table = {l.split()[0]: l.split()[1:] for l in open(os.path.join(os.path.dirname(cd.__file__), "data",
                                                                    "atomic_scattering_factors.txt"))
             if l.strip() and not l.startswith("#")}
    a = [float(v) for v in table["O"]]
    assert [float(v) for v in groups[1][4:]] == pytest.approx(a[0:5] + a[6:11] + [a[5]], abs=5e-7)


def test_xyz_errors(tmp_path):
    bad = tmp_path / "bad.xyz"
    bad.write_text("Xx 0 0 0\n")
    with pytest.raises(cd.DebyeError) as e:
        cd.xyz_to_deb(bad)
    assert e.value.code == cd.DBX_ERR_PARSE
    ragged = tmp_path / "ragged.xyz"
    ragged.write_text("O 0 0 0 1 0\nO 1 1 1\n")
    with pytest.raises(cd.DebyeError):
        cd.xyz_to_deb(ragged)
    with pytest.raises(cd.DebyeError) as e:
        cd.xyz_to_deb(tmp_path / "missing.xyz")
    assert e.value.code == cd.DBX_ERR_IO


def _debi(path):
    d = np.loadtxt(path)
    return d[:, 0], d[:, 1]


@pytest.mark.parametrize("norm", ["none", "int", "max"])
def test_average_debi_against_scipy(golden_dir, tmp_path, norm):
    """Three patterns over different 2theta ranges: the common range at 0.02 deg, each raw pattern
    through its not-a-knot cubic spline, summed, normalised."""
    t1, y1 = _debi(golden_dir / "UH3_7nm.debI.ref")
    t2, y2 = _debi(golden_dir / "GammaAl2O3_10nm.debI.ref")
    p3 = tmp_path / "c.debI"
    t3 = np.linspace(20.0, 120.0, 501)
    y3 = 1000 + 500 * np.sin(t3 / 7.0) + 0.1 * t3 ** 2
    np.savetxt(p3, np.c_[t3, y3], fmt="%.10g", delimiter="\t")
    paths = [golden_dir / "UH3_7nm.debI.ref", golden_dir / "GammaAl2O3_10nm.debI.ref", p3]
    tt, I = cd.average_debi(paths, norm=norm)
    lo = max(0.0, t1.min(), t2.min(), t3.min())
    hi = min(180.0, t1.max(), t2.max(), t3.max())
    n = int(np.floor((hi - lo) / 0.02 + 1e-9)) + 1
    grid = lo + 0.02 * np.arange(n)
    acc = sum(CubicSpline(t, y, bc_type="not-a-knot")(grid) for t, y in ((t1, y1), (t2, y2), (t3, y3)))
    ref = {"none": acc / 3, "int": acc / np.trapezoid(acc, grid), "max": acc / acc.max()}[norm]
    assert np.array_equal(tt, grid)
    assert np.max(np.abs(I - ref) / np.abs(ref)) < 1e-11


def test_average_cli_and_spline_small_cases(tmp_path):
    """Four and five knots (not-a-knot: one cubic through four points), exact on cubics."""
    for n in (4, 5, 9):
        t = np.sort(np.random.default_rng(n).uniform(0, 50, n))
        y = 2 + 0.5 * t - 0.03 * t ** 2 + 0.0007 * t ** 3
        p = tmp_path / f"p{n}.debI"
        np.savetxt(p, np.c_[t, y], fmt="%.17g", delimiter="\t")
        tt, I = cd.average_debi([p], step=0.5)
        assert np.allclose(I, 2 + 0.5 * tt - 0.03 * tt ** 2 + 0.0007 * tt ** 3, rtol=1e-10, atol=1e-9)
    out = tmp_path / "avg.debI"
    subprocess.run([CLI, "--average", "--norm", "max", "-o", str(out), str(tmp_path / "p9.debI")], check=True)
    d = np.loadtxt(out)
    assert abs(d[:, 1].max() - 1.0) < 1e-9


# ---- structure view (Matlab/viewdeb.m) ------------------------------------------------------
def _radii_table():
    rows = {}
    path = os.path.join(os.path.dirname(cd.__file__), "data", "atomic_radii.txt")
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        s, r, *rgb = line.split()
        rows[s.lower()] = (float(r), tuple(float(v) for v in rgb))
    return rows


def _svg(path):
    import xml.etree.ElementTree as ET
    ns = "{http://www.w3.org/2000/svg}"
    root = ET.parse(path).getroot()
    circles = [(float(c.get("cx")), float(c.get("cy")), float(c.get("r")), c.get("fill")) for c in root.iter(ns + "circle")]
    texts = [(t.text, t.get("fill")) for t in root.iter(ns + "text") if t.get("font-size") == "24"]
    return circles, texts


def _hex(rgb):
    return "#" + "".join(f"{int(round(min(1, max(0, v)) * 255)):02x}" for v in rgb)


def test_view_deb_clips_and_draws_the_reference_example(golden_dir, tmp_path):
    """Data/UH3_7nm.deb with the default limits (-10 < coordinate < 10 on every axis): the kept
    atoms are exactly those of the file inside the box, listed in the .xyz in file order with
    their elements; the SVG draws one disc per atom, of radius 5 r (r from the radius table) in
    the element's colour, and lists H and U in their colours."""
    deb = golden_dir / "UH3_7nm.deb"
    head, groups, pos = _deb_groups(deb)
    n = [int(g[0]) for g in groups]
    elem = np.repeat([g[1] for g in groups], n)
    p32 = pos.astype(np.float32)
    keep = np.all((p32 > -10) & (p32 < 10), axis=1)
    svg, xyz = tmp_path / "v.svg", tmp_path / "v.xyz"
    assert cd.view_deb(deb, svg=svg, xyz=xyz) == int(keep.sum()) > 0
    lines = open(xyz).read().splitlines()
    assert int(lines[0]) == keep.sum() and lines[1] == "UH3_7nm.deb"
    got = [l.split() for l in lines[2:]]
    assert [g[0] for g in got] == list(elem[keep])
    assert np.allclose(np.array([[float(v) for v in g[1:]] for g in got]), p32[keep], rtol=1e-6, atol=1e-6)
    table = _radii_table()
    circles, texts = _svg(svg)
    assert len(circles) == keep.sum()
    want = {e: (5 * table[e.lower()][0], _hex(table[e.lower()][1])) for e in ("H", "U")}
    assert sorted(set((round(r, 3), f) for _, _, r, f in circles)) == sorted(set((round(r, 3), f) for r, f in want.values()))
    assert texts == [("H", want["H"][1]), ("U", want["U"][1])]


def test_view_deb_limits_types_and_projection(tmp_path):
    """Per-axis limits (open intervals), subgroup types with charge and label suffixes mapped to
    their elements, an element missing from the table kept in the .xyz but not drawn, and the
    view's vertical z axis: atoms that differ in z alone are drawn one above the other, higher z
    higher on the page."""
    types = ["Al3+", "O2-", "Fe:oct", "Cu1", "Xx"]
    pts = [[(0, 0, -1), (0, 0, 1), (0, 0, 2.5)], [(1, 2, 3)], [(5, 0, 0), (-4, 0, 0)], [(0, 3.9, 0)], [(1, 1, 1)]]
    deb = tmp_path / "m.deb"
    with open(deb, "w") as f:
        f.write("0 16 2048\n1.5418 10 140 0.02\n20 0.001 0\n%d\n" % len(types))
        for t, p in zip(types, pts):
            f.write(f"{len(p)} {t} 1 0 " + " ".join(["1"] * 11) + "\n")
        for p in pts:
            for v in p:
                f.write("%g\t%g\t%g\n" % v)
    xyz, svg = tmp_path / "m.xyz", tmp_path / "m.svg"
    # x in (-4.5, 4.5), y in (-10, 3.9) (3.9 itself excluded), z in (-2, 2.5) (2.5 excluded)
    n = cd.view_deb(deb, limits=((-4.5, 4.5), (-10, 3.9), (-2, 2.5)), svg=svg, xyz=xyz)
    rows = [l.split() for l in open(xyz).read().splitlines()[2:]]
    assert n == 4 and [r[0] for r in rows] == ["Al", "Al", "Fe", "Xx"]
    circles, texts = _svg(svg)
    assert [t for t, _ in texts] == ["Al", "O", "Fe", "Cu", "Xx"]
    assert len(circles) == 3  # Xx: no radius, not drawn
    # the two Al atoms (z = -1 and z = 1) in drawing order, farthest first: from above the xy
    # plane the z = 1 atom is the nearer one, drawn second, straight above the other (smaller cy)
    al = [c for c in circles if c[3] == _hex(_radii_table()["al"][1])]
    assert len(al) == 2 and abs(al[0][0] - al[1][0]) < 1e-6 and al[1][1] < al[0][1]
    assert [float(r[3]) for r in rows[:2]] == [-1.0, 1.0]  # the .xyz keeps the file's order


def test_cli_view(golden_dir, tmp_path):
    deb = tmp_path / "UH3_7nm.deb"
    shutil.copy(golden_dir / "UH3_7nm.deb", deb)
    r = subprocess.run([CLI, "--view", str(deb), "--limits", "-5", "5"], check=True, capture_output=True, text=True)
    n = int(r.stdout.split()[0])
    circles, _ = _svg(str(deb) + ".svg")
    assert n == len(circles) == cd.view_deb(deb, limits=(-5, 5), xyz=tmp_path / "a.xyz")
