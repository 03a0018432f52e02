"""CPU checks of the simulation driver's argument handling (rbc_run_main): the reference
parser's errors for missing mandatory arguments and values this build does not accept end the
run with status 1 before any GPU work."""
import ctypes

import pytest


@pytest.fixture(scope="module")
def lib():
    from sdc_amd import rbc
    L = rbc.load_library()
    L.rbc_run_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    return L


def _call(lib, args):
    argv = (ctypes.c_char_p * (len(args) + 1))(*([b"rbc_run"] + [a.encode() for a in args]))
    return lib.rbc_run_main(len(args) + 1, argv)


BASE = ["-nmodes", "32", "32", "8", "-clength", "4", "4", "-ra", "5000", "-pr", "1", "-dt", "0.001", "-t", "0.01"]


@pytest.mark.parametrize("args", [
    [],                                            # -nmodes missing
    BASE[:4] + BASE[6:],                           # -clength missing (and -nmodes incomplete)
    [a for a in BASE if a not in ("-ra", "5000")],  # -ra missing
    BASE[:-4] + ["-t", "0.01"],                    # -dt missing
    BASE + ["-precision", "f16"],                  # unknown precision
    BASE + ["-tracer_mode", "euler"],              # unknown tracer integrator
    BASE + ["-nfiles", "x"],                       # value that does not parse
])
def test_driver_rejects(lib, args):
    assert _call(lib, args) == 1
