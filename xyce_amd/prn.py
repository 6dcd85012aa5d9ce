"""Waveform files in the layout of Xyce's .PRINT TRAN output (FORMAT=STD,
the .prn file; N_IO_Outputter.C TimePrn and the column set-up of
fixupColumns / printHeader / printValue), so that results can be diffed
against any reference run:

  * columns INDEX (left-justified, fixed, width 5, shown as "Index"), TIME,
    then one per printed quantity (scientific, precision 9, width 17,
    names centred in the header, values right-justified), separated by
    one space;
  * one row per output point, then "End of Xyce(TM) Simulation".
"""
import numpy as np

WIDTH, PRECISION = 17, 9


def _centre(name, width=WIDTH):
    if width <= len(name):
        return name
    left = (width - len(name)) // 2
    return " " * left + name + " " * (width - left - len(name))


def write_prn(path, times, names, waves):
    """times: (n,), names: quantity names such as "V(OUT0)", waves:
    (len(names), n)."""
    waves = np.asarray(waves, dtype=np.float64)
    times = np.asarray(times, dtype=np.float64)
    if waves.shape != (len(names), len(times)):
        raise ValueError("waves must be (len(names), len(times))")
    with open(path, "w") as f:
        f.write(" ".join(["Index", _centre("TIME")] + [_centre(n) for n in names]) + "\n")
        for k, t in enumerate(times):
            row = ["%-5d" % k, "%*.*e" % (WIDTH, PRECISION, t)]
            row += ["%*.*e" % (WIDTH, PRECISION, v) for v in waves[:, k]]
            f.write(" ".join(row) + "\n")
        f.write("End of Xyce(TM) Simulation\n")


def read_prn(path):
    """Returns (names, times, waves) of a .prn file in this layout."""
    with open(path) as f:
        lines = f.read().splitlines()
    head = lines[0].split()
    if head[:2] != ["Index", "TIME"]:
        raise ValueError("not a TIME .prn file")
    rows = []
    for ln in lines[1:]:
        if ln.startswith("End of Xyce"):
            break
        rows.append([float(v) for v in ln.split()])
    data = np.array(rows).reshape(-1, len(head))
    return head[2:], data[:, 1], data[:, 2:].T


def probe_names(nodes):
    """Xyce's column names of node voltages: V(NAME), upper case."""
    return ["V(%s)" % n.upper() for n in nodes]
