"""GPU parity of the BSIM3 charge models capMod 0, 1 and 2
(N_DEV_MOSFET_B3.C:6005-6685, overlaps 7164-7215) against the oracle, in
every solver state of a transient and in each eval kernel form, on the
variety netlist (its cards cover the 0/100, 40/60 and 50/50 partitions,
both polarities, sheet resistance and binning)."""
import pytest

from test_gpu_bsim3 import SOLVER_STATES, check, run_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("capmod", [0, 1, 2])
@pytest.mark.parametrize("si", range(len(SOLVER_STATES)))
def test_bsim3_capmod_parity(capmod, si):
    import torch
    from xyce_amd import circuits
    nl = circuits.bsim3_variety(seed=11 + capmod, capmod=capmod)
    check(*run_pair(nl, SOLVER_STATES[si], 20 + si, torch))


@pytest.mark.parametrize("capmod", [0, 1, 2])
@pytest.mark.parametrize("variant", ["hbm", "staged", "staged_split"])
def test_bsim3_capmod_every_kernel_form(capmod, variant):
    import torch
    from xyce_amd import circuits
    nl = circuits.bsim3_variety(seed=31, capmod=capmod)
    check(*run_pair(nl, SOLVER_STATES[5], 7, torch, variant=variant))


def test_bsim3_mixed_capmods_in_one_circuit():
    """Cards of different charge models side by side in one launch."""
    import torch
    from xyce_amd import circuits
    nl = circuits.bsim3_variety(seed=5)
    for k, name in enumerate(nl.models):
        nl.models[name].params["CAPMOD"] = float(k % 4)
        nl.models[name].params["VFBCV"] = -0.9
    check(*run_pair(nl, SOLVER_STATES[3], 3, torch))
    check(*run_pair(nl, SOLVER_STATES[5], 4, torch))
