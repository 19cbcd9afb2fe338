"""Host-side behaviour of the C++ drop-in that needs no GPU: module import and
meta::Inputparameters option handling (reference option names, defaults,
rejection of unknown options as boost::program_options does)."""
import pytest

from cl2qcd_amd import _host as H


def test_module_exports_the_interface():
    for name in ("Lattice", "Gaugefield", "Spinorfield", "Spinorfield_eo", "Staggeredfield_eo", "Aee",
                 "Aee_AND_gamma5_eo", "Aee_minus", "Aee_minus_AND_gamma5_eo", "Qplus_eo", "Qminus_eo",
                 "QplusQminus_eo", "D_KS_eo", "MdagM_eo", "cg", "bicgstab", "cg_m", "dslash", "DKS_eo",
                 "SolverStuck", "SolverDidNotSolve"):
        assert hasattr(H, name), name
    assert issubclass(H.SolverStuck, H.SolverException) and issubclass(H.SolverDidNotSolve, H.SolverException)


@pytest.mark.parametrize("bad", [["--nSpace=5"], ["--bogus=1"], ["--fermionAction=clover_x"], ["--solver=gmres"],
                                 ["--kappa=abc"], ["nSpace=4"], ["--nDevices=2"], ["--precision=32"]])
def test_rejects_bad_options(bad):
    with pytest.raises(ValueError):
        H.Lattice(bad)
