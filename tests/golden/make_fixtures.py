"""Write the golden fixtures under tests/golden/ from data held in the
reference tree (run in a container where /root/reference exists):

    python tests/golden/make_fixtures.py [/root/reference]

Only numbers are extracted (positions, coefficients, histogram counts and the
expected values of the reference's own tests); no source text is copied.
"""
from __future__ import annotations

import json
import pathlib
import re
import sys

HERE = pathlib.Path(__file__).resolve().parent


def _floats(text: str):
    return [float(x) for x in re.findall(r"-?\d+\.\d+(?:[eE][-+]?\d+)?|-?\d+(?:[eE][-+]?\d+)", text)]


def _split_top(body: str):
    depth = 0
    args, cur = [], ""
    for ch in body:
        if ch in "[(":
            depth += 1
        if ch in "])":
            depth -= 1
            if depth < 0:
                break
        if ch == "," and depth == 0:
            args.append(cur.strip())
            cur = ""
        else:
            cur += ch
    args.append(cur.strip())
    return args


def ne13_setup(ref: pathlib.Path):
    src = (ref / "scripts" / "ne13.jl").read_text()
    c_block = src[src.index("c = ["): src.index("]", src.index("c = ["))]
    coeff = _floats(c_block)
    p_block = src[src.index("pos_ne13 = ["): src.index("\n]", src.index("pos_ne13 = ["))]
    pos = _floats(p_block)
    ex = (ref / "examples" / "ne13.jl").read_text()
    a2b = float(re.search(r"AtoBohr\s*=\s*([0-9.]+)", ex).group(1))
    rad = float(re.search(r"radius\s*=\s*([0-9.]+)\s*\*\s*AtoBohr", ex).group(1))
    out = dict(
        source="scripts/ne13.jl:39-46 (coefficients), :60-74 (positions, Angstrom); examples/ne13.jl:56,66",
        coeff=coeff,
        pos_angstrom=[pos[i: i + 3] for i in range(0, len(pos), 3)],
        AtoBohr=a2b,
        radius_angstrom=rad,
    )
    assert len(out["pos_angstrom"]) == 13 and len(coeff) == 6
    (HERE / "ne13_setup.json").write_text(json.dumps(out, indent=1))


def multihist_ne13(ref: pathlib.Path):
    s = (ref / "test" / "multihist_test.jl").read_text()
    i = s.index("results=Output{Float64}(")
    line = s[i: s.index("\n", i)]
    args = _split_top(line[len("results=Output{Float64}("):])
    hist = [_floats(h) for h in re.findall(r"\[([^\[\]]*)\]", args[8])]
    temp = re.search(r"TempGrid\{(\d+)\}\((\d+),\s*(\d+)\)", s[: s.index("results=")])
    cv = {int(k): float(v) for k, v in re.findall(r"cv\[(\d+)\]==([-0-9.eE]+)", s)}
    out = dict(
        source="test/multihist_test.jl:5-15 (reference Ne13 NVT Output and expected multihistogram_NVT values)",
        n_temps=int(temp.group(1)),
        t_min=float(temp.group(2)),
        t_max=float(temp.group(3)),
        n_bin=int(args[0]),
        en_min=float(args[1]),
        en_max=float(args[2]),
        en_avg=_floats(args[6]),
        heat_cap=_floats(args[7]),
        en_histogram=hist,
        count_stat_atom=_floats(args[11]),
        count_stat_exc=_floats(args[14]),
        conv_threshold=1e-3,
        expected_cv=cv,
        expected_len=int(re.search(r"length\(cv\)==(\d+)", s).group(1)),
    )
    assert len(hist) == out["n_temps"] and all(len(h) == out["n_bin"] + 2 for h in hist)
    (HERE / "multihist_ne13.json").write_text(json.dumps(out))


def multihist_npt32(ref: pathlib.Path):
    """The NPT testset of test/multihist_test.jl:19-31: a 3-temperature run
    of 32 atoms (20-40 K) with its (energy, volume) histograms and four
    expected heat capacities.  The testset sits inside a comment block in the
    reference and its expected values index the heat-capacity array; the
    reference function returns (temperatures, heat capacities).  The
    histograms are stored sparse: [temperature, energy row, volume column,
    count] with rows and columns 0..n_bin+1 (underflow, bins, overflow)."""
    s = (ref / "test" / "multihist_test.jl").read_text()
    s = s[s.index('@testset "multihist_NPT"'):]
    ens = re.search(r"NPT\((\d+),\s*([-0-9.eE]+)", s)
    temp = re.search(r"TempGrid\{(\d+)\}\((\d+),\s*(\d+)\)", s)
    i = s.index("Output{Float64}(")
    line = s[i: s.index("\n", i)]
    args = _split_top(line[len("Output{Float64}("):])
    mats = re.findall(r"\[([^\[\]]*)\]", args[9])
    n_temps, n_bin = int(temp.group(1)), int(args[0])
    sparse = []
    for t, m in enumerate(mats):
        rows = [r.split() for r in m.split(";")]
        assert len(rows) == n_bin + 2 and all(len(r) == n_bin + 2 for r in rows)
        for e, r in enumerate(rows):
            for v, x in enumerate(r):
                if float(x) != 0.0:
                    sparse.append([t, e, v, float(x)])
    cp = {int(k): float(v) for k, v in re.findall(r"cp_3\[(\d+)\]==([-0-9.eE]+)", s)}
    out = dict(
        source="test/multihist_test.jl:19-31 (reference NPT Output of 32 atoms and expected multihistogram_NPT values)",
        n_atoms=int(ens.group(1)),
        pressure=float(ens.group(2)),
        n_temps=n_temps,
        t_min=float(temp.group(2)),
        t_max=float(temp.group(3)),
        n_bin=n_bin,
        en_min=float(args[1]),
        en_max=float(args[2]),
        v_min=float(args[3]),
        v_max=float(args[4]),
        en_avg=_floats(args[6]),
        heat_cap=_floats(args[7]),
        ev_histogram_sparse=sparse,
        conv_threshold=1e-3,
        expected_cp=cp,
        expected_len=int(re.search(r"length\(cp_3\)==(\d+)", s).group(1)),
    )
    assert len(mats) == n_temps and len(cp) == 4
    (HERE / "multihist_npt32.json").write_text(json.dumps(out))


def eljb_pbc216(ref: pathlib.Path):
    """ELJPotentialB coefficients of the reference's periodic neon scripts
    (scripts/pbc216.jl:56-59: a, b and c1 of ELJPotentialB{6}(a, b, c1))."""
    s = (ref / "scripts" / "pbc216.jl").read_text()
    vec = lambda name: _floats(re.search(r"\n" + name + r" = \[([^\]]*)\]", s).group(1))
    out = dict(source="scripts/pbc216.jl:56-59", a=vec("a"), b=vec("b"), c=vec("c1"))
    assert len(out["a"]) == len(out["b"]) == len(out["c"]) == 6
    (HERE / "eljb_pbc216.json").write_text(json.dumps(out, indent=1))


def nnp2_potentialfile(ref: pathlib.Path):
    """The two-species RuNNer potential of the reference's tests
    (test/potentialfile.jl: radial X and angular V tables, network shape,
    species counts) and its known answer (test/test_potentials.jl:68-88:
    six atoms, E = -0.261652899); the second species' weights
    (test/testing_data/weights.030.data) are copied next to weights.029."""
    pf = (ref / "test" / "potentialfile.jl").read_text()
    xb = pf[pf.index("X = ["): pf.index("]", pf.index("X = ["))]
    X = [[float(v) for v in line.split()] for line in xb.splitlines()[1:] if line.strip()]
    vb = pf[pf.index("V = ["): pf.index("\n]", pf.index("V = ["))]
    V = [[float(v) for v in m.split(",")] for m in re.findall(r"\[([^\[\]]*)\]", vb[len("V = ["):])]
    tp = (ref / "test" / "test_potentials.jl").read_text()
    tb = tp[tp.index('@testset "RuNNerPotentialTest"'):]
    pos = [[float(v) for v in m.split(",")] for m in re.findall(r"v\d = SVector\(([^)]*)\)", tb)]
    pos = [[p[0], p[1], p[2]] for p in pos]
    # "5.0 * 3^0.5"-style entries do not occur here: every coordinate is a literal
    assert len(pos) == 6 and all(len(p) == 3 for p in pos)
    out = dict(
        source="test/potentialfile.jl:6-129 (tables), test/test_potentials.jl:68-88 (configuration, energy)",
        radial=X, angular=V,
        num_nodes=[int(v) for v in re.search(r"num_nodes::Vector\{Int32\} = \[([^\]]*)\]", pf).group(1).split(",")],
        activations=[int(v) for v in re.search(r"activation_functions::Vector\{Int32\} = \[([^\]]*)\]",
                                               pf).group(1).split(",")],
        n1=4, n2=2,
        positions=pos,
        radius=float(re.search(r"radius=([0-9.]+)", tb).group(1)),
        energy=float(re.search(r"E ≈ (-?[0-9.]+)", tb).group(1)),
    )
    assert len(X) == 5 and len(V) == 26 and out["num_nodes"] == [88, 20, 20, 1]
    assert all(len(r) == 4 for r in X) and all(len(r) == 4 for r in V)
    (HERE / "nnp2_brass.json").write_text(json.dumps(out, indent=1))
    src = ref / "test" / "testing_data" / "weights.030.data"
    (HERE / "nnp" / "weights.030.data").write_bytes(src.read_bytes())


def nnvt_states(ref: pathlib.Path):
    """The NNVT testset of test/test_states.jl:110-182: six atoms (four of
    species 1, two of species 2) in a sphere of radius 8, the potential of
    test/potentialfile.jl, and the energy change of swapping atoms 3 and 6
    (1-based; the indices follow from the test's own assertions on the
    swapped positions, :179-180), delen2 = -0.0017084901948."""
    s = (ref / "test" / "test_states.jl").read_text()
    s = s[s.index('@testset "States NNVT"'):]
    s = s[: s.index("\nend")]
    pos = [[float(v) for v in m.split(",")] for m in re.findall(r"v\d = SVector\(([^)]*)\)", s)]
    swap = re.search(r"state\.config\[(\d+)\] == refstate\.config\[(\d+)\]", s)
    out = dict(
        source="test/test_states.jl:110-182 (NNVT configuration, swap indices, swap energy change)",
        positions=pos,
        n1=int(re.search(r"NNVT\(\[(\d+), (\d+)\]\)", s).group(1)),
        n2=int(re.search(r"NNVT\(\[(\d+), (\d+)\]\)", s).group(2)),
        radius=float(re.search(r"radius=([0-9.]+)", s).group(1)),
        swap=[int(swap.group(1)), int(swap.group(2))],
        delta_swap=float(re.search(r"delen2 ≈ (-?[0-9.eE-]+)", s).group(1)),
        delta_move_rng=float(re.search(r"delen ≈ (-?[0-9.eE-]+)", s).group(1)),
    )
    assert len(pos) == 6 and out["n1"] == 4 and out["n2"] == 2 and out["swap"] == [3, 6]
    (HERE / "nnvt_states.json").write_text(json.dumps(out, indent=1))


def known_answers(ref: pathlib.Path):
    """Deterministic known answers of the reference's tests, with the lines
    they come from.  Each value is read from the file it cites."""
    tp = (ref / "test" / "test_potentials.jl").read_text()
    rt = (ref / "test" / "runtests.jl").read_text()
    eam = float(re.search(r"E ≈ (-?[0-9.]+)", tp).group(1))
    out = dict(
        eam_three_atoms=dict(source="test/test_potentials.jl:44-56", positions=[[1.0, 2.0, 3.0], [2.0, 4.0, 6.0],
                                                                                [0.0, 1.0, 0.0]],
                             n=1.0, m=1.0, eps=1.0, C=1.0, a=1.0, ean=0.5, energy=eam),
        elj=dict(source="test/test_potentials.jl:3-15",
                 cases=[dict(kind="elj", coeff=[-2.0, 0.0, 1.0], r2=1.0, energy=-1.0),
                        dict(kind="elj_even", coeff=[-2.0, 1.0], r2=1.0, energy=-1.0),
                        dict(kind="elj", coeff=[-1.0, 2.0, 3.0, -4.0, 5.0], r2=1.0, energy=5.0)]),
        distance=dict(source="test/runtests.jl:74-100",
                      d2mat_13=float(re.search(r"d2mat\[1, 3\] == ([0-9.]+)", rt).group(1)),
                      v1=[1.0, 2.0, 3.0], v2=[2.0, 4.0, 6.0], v3_sphere=[0.0, 1.0, 0.0],
                      v3_cubic=[3.0, 6.0, 9.0], box=10.0, cubic_13=36.0, plain_13=56.0, plain_12=14.0),
        volume=dict(source="test/runtests.jl:198-202", box=10.0, volume=1000.0),
        tempgrid=dict(source="test/runtests.jl:216-235", kB=3.16681196e-6),
        checkpoint_r2=dict(source="test/testing_data/checkpoint/config.1:2",
                           r2=float(re.search(r"r2:\s*([0-9.]+)", (ref / "test/testing_data/checkpoint/config.1")
                                              .read_text()).group(1))),
    )
    (HERE / "known_answers.json").write_text(json.dumps(out, indent=1))


def cu78_start(ref: pathlib.Path):
    src = (ref / "scripts" / "run_Cu78.jl").read_text()
    blk = src[src.index("pos_cu78 = ["): src.index("\n]", src.index("pos_cu78 = ["))]
    pos = _floats(blk)
    rad = float(re.search(r"SphericalBC\(;\s*radius\s*=\s*([0-9.]+)\s*\*\s*AtoBohr", src).group(1))
    t = re.search(r"ti\s*=\s*([0-9.]+)\s*\n\s*tf\s*=\s*([0-9.]+)", src)
    n_traj = int(re.search(r"n_traj\s*=\s*(\d+)", src).group(1))
    out = dict(
        source="scripts/run_Cu78.jl:48-127 (positions, Angstrom, centred at :140-149), :155 (container), "
               ":17-21 (temperatures)",
        pos_angstrom=[pos[i: i + 3] for i in range(0, len(pos), 3)],
        radius_angstrom=rad,
        t_range=[float(t.group(1)), float(t.group(2))],
        n_traj=n_traj,
    )
    assert len(out["pos_angstrom"]) == 78
    (HERE / "cu78_start.json").write_text(json.dumps(out, indent=1))


def ne38_start(ref: pathlib.Path):
    src = (ref / "scripts" / "ne38_B.jl").read_text()
    blk = src[src.index("pos_ne38 = ["): src.index("\n]", src.index("pos_ne38 = ["))]
    pos = _floats(blk)
    rad = float(re.search(r"SphericalBC\(;\s*radius\s*=\s*([0-9.]+)\)", src).group(1))
    t = re.search(r"ti\s*=\s*([0-9.]+)\s*\n\s*tf\s*=\s*([0-9.]+)", src)
    n_traj = int(re.search(r"n_traj\s*=\s*(\d+)", src).group(1))
    out = dict(
        source="scripts/ne38_B.jl:48-87 (positions, Angstrom), :97 (container, Angstrom), :12-14 (temperatures)",
        pos_angstrom=[pos[i: i + 3] for i in range(0, len(pos), 3)],
        radius_angstrom=rad,
        t_range=[float(t.group(1)), float(t.group(2))],
        n_traj=n_traj,
    )
    assert len(out["pos_angstrom"]) == 38
    (HERE / "ne38_start.json").write_text(json.dumps(out, indent=1))


def main():
    ref = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    ne13_setup(ref)
    multihist_ne13(ref)
    multihist_npt32(ref)
    eljb_pbc216(ref)
    nnp2_potentialfile(ref)
    nnvt_states(ref)
    known_answers(ref)
    cu78_start(ref)
    ne38_start(ref)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
