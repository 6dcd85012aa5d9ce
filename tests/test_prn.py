"""The .prn writer: the layout of Xyce's FORMAT=STD .PRINT TRAN file
(N_IO_Outputter.C printHeader / printValue / TimePrn::doFinishOutput),
checked character for character on a small case, and a round trip."""
import numpy as np

from xyce_amd.prn import probe_names, read_prn, write_prn


def test_layout_is_xyce_std(tmp_path):
    p = tmp_path / "a.prn"
    write_prn(p, [0.0, 1e-12], ["V(OUT)"], [[1.8, -2.5e-3]])
    lines = p.read_text().splitlines()
    # "Index" (width 5), then TIME and V(OUT) centred in 17 columns
    assert lines[0] == "Index" + " " + " " * 6 + "TIME" + " " * 7 + " " + " " * 5 + "V(OUT)" + \
        " " * 6
    # one space before every column after the first, values right-justified
    assert lines[1] == "0    " + " " + "  0.000000000e+00" + " " + "  1.800000000e+00"
    assert lines[2] == "1    " + " " + "  1.000000000e-12" + " " + " -2.500000000e-03"
    assert lines[3] == "End of Xyce(TM) Simulation"


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    t = np.linspace(0, 2e-9, 41)
    w = rng.uniform(-2, 2, (3, 41))
    names = probe_names(["out0", "n5", "vdd"])
    assert names == ["V(OUT0)", "V(N5)", "V(VDD)"]
    write_prn(tmp_path / "b.prn", t, names, w)
    n2, t2, w2 = read_prn(tmp_path / "b.prn")
    assert n2 == names
    assert np.allclose(t2, t, rtol=1e-9, atol=0)
    assert np.allclose(w2, w, rtol=1e-9, atol=0)
