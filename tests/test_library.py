"""CPU checks of the C ABI library: it loads, exports every entry point the header declares, fills
the reference defaults, and fails loudly (no CPU fallback) when no GPU is visible."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from polystokes_brownian_amd import _abi
from polystokes_brownian_amd import polystokes as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "polystokes_gpu.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ps_[a-z_0-9]+)\s*\(", src)))


def test_header_symbols_exported():
    names = declared()
    assert "ps_saddle_solve" in names and "ps_sqrt_apply" in names and "ps_mob_mult" in names
    out = subprocess.run(["nm", "-D", "--defined-only", P.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (ps_[a-z_0-9]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = P.lib()
    for n in names:
        assert getattr(lib, n) is not None


def test_struct_sizes_match_header():
    # ctypes mirror of ps_params / ps_solve_info has the C layout (checked by a tiny C program)
    code = r'''
    #include <stdio.h>
    #include <stddef.h>
    #include "polystokes_gpu.h"
    int main(){ printf("%zu %zu %zu %zu %zu %zu\n", sizeof(ps_params), sizeof(ps_solve_info), sizeof(ps_sqrt_info),
                       sizeof(ps_counters), sizeof(ps_profile), offsetof(ps_params, ir_maxit)); return 0; }'''
    tmp = os.path.join(os.environ.get("TMPDIR", "/tmp"), "ps_sizes_check")
    with open(tmp + ".c", "w") as fh:
        fh.write(code)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", tmp, tmp + ".c"], check=True)
    got = list(map(int, subprocess.run([tmp], capture_output=True, text=True, check=True).stdout.split()))
    want = [C.sizeof(_abi.PsParams), C.sizeof(_abi.PsSolveInfo), C.sizeof(_abi.PsSqrtInfo),
            C.sizeof(_abi.PsCounters), C.sizeof(_abi.PsProfile), _abi.PsParams.ir_maxit.offset]
    assert got == want


def test_reference_defaults():
    p = P.default_params()
    # src/bindings.cpp:56-69 and src/init.cpp:64
    assert (p.mm_HI, p.chain_HI, p.fene, p.record_forces, p.tether, p.mono_ev) == (1, 0, 1, 1, 0, 0)
    assert p.seed == 12345 and p.restart == 0.0 and p.box_active == 0
    assert p.ksp_rtol == 1e-6 and p.ksp_abstol == 1e-50 and p.ksp_maxit == 10000
    assert p.weaken_trap == -1 and p.hs_kpref == 1.0 and p.hs_exp_k == 100.0


def _gpu_visible():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly():
    p = P.default_params(Np=10, Nm=8, Nc=2, Npoly=2, Nmono_per_chain=4, beta=0.2, dt=1e-3)
    with pytest.raises(P.PsError) as ei:
        P.Context(p)
    assert ei.value.code == _abi.PS_ERR_NODEV


def test_argument_errors_before_device():
    p = P.default_params(Np=10, Nm=7, Nc=2, beta=0.2, dt=1e-3)
    with pytest.raises(P.PsError) as ei:
        P.Context(p)
    assert ei.value.code == _abi.PS_ERR_ARG
    p = P.default_params(Np=10, Nm=8, Nc=2, beta=0.2, dt=1e-3, mm_HI=0)
    with pytest.raises(P.PsError) as ei:
        P.Context(p)
    assert "mm_HI" in str(ei.value)


def test_polystokes_api_surface():
    sim = P.PolyStokes(1e-3, 10, 0.1, "out", mm_HI=True, chain_HI=True, box=[10, 10, 10], seed=7)
    assert sim.nsteps == 100
    sim.particle_info(1.0, 1.0, 12, 2, 10, 2, 5, 0.2, 100.0, 0.42, 0.6, 1000)
    sim.trap_info(500.0, 0.0, 1.0)
    b = P.Box()
    b.configure([10.0, 10.0, 10.0])
    assert b.active() and b.minimum_image(6.0, -6.0, 4.0) == (-4.0, 4.0, 4.0)
    with pytest.raises(ValueError):
        b.configure([1.0, 2.0])


def test_frame_file_time_format():
    """Frame names use a default std::ostream's formatting of t (src/write_configuration.cpp:18-19),
    i.e. C's %g with 6 significant digits."""
    from polystokes_brownian_amd.frames import format_time, frame_path
    cases = {0.001: "0.001", 1e-05: "1e-05", 0.123456789: "0.123457", 100.0: "100", 1234567.0: "1.23457e+06",
             0.0025: "0.0025", 0.0: "0", 8e-06: "8e-06", 3 * 8e-06: "2.4e-05", 0.1 + 0.2: "0.3"}
    for t, want in cases.items():
        assert format_time(t) == want, t
        assert format_time(t) == "%g" % t
    assert frame_path("out", "config_", 0.001) == os.path.join("out", "config_0.001_.txt")


@pytest.mark.parametrize("chunk", [0, 1, 7, 4096, 60001, 250000])
def test_normal_stream_is_std_normal_distribution(chunk):
    """The library's bulk noise generator reproduces std::mt19937_64 + std::normal_distribution bit for
    bit (the oracle calls the standard library itself), across fill boundaries where the polar
    method's cached second value carries over, and with the transform split over host threads."""
    import ctypes as C
    from oracle.pyoracle import normals
    from polystokes_brownian_amd.polystokes import lib
    L = lib()
    L.ps_normal_stream.argtypes = [C.c_uint64, C.POINTER(C.c_double), C.c_int64, C.c_int64, C.c_double]
    n = 600001
    for seed in (12345, 2 ** 63 + 17):
        want = normals(seed, n)
        # margin 0.3: each fill first takes too few attempts and extends over several further ranges
        for margin in (0.0, 0.3):
            got = np.zeros(n)
            assert L.ps_normal_stream(seed, got.ctypes.data_as(C.POINTER(C.c_double)), n, chunk, margin) == 0
            np.testing.assert_array_equal(got, want)
