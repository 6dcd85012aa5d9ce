"""Checkpoint files refuse a circuit of other sizes (host-side check, no
GPU needed: the checkpoint records the sizes it was made for)."""
import numpy as np
import pytest


def test_checkpoint_size_guard(tmp_path):
    from xyce_amd.transient import Transient

    class Fake:
        _CKPT = Transient._CKPT
        restore = Transient.restore

    f = Fake()
    f.dev = type("D", (), {"nnz": 10})()
    f.x = np.zeros(5)
    f.sta_c = np.zeros(2)
    f.sto_c = np.zeros(3)
    np.savez(tmp_path / "c.npz", sizes=np.array([11, 5, 2, 3]))
    with pytest.raises(ValueError):
        f.restore(tmp_path / "c.npz")
