"""The C ABI (CPU only): libsmectic.so loads and exports every entry point that
include/smectic.h declares, and the ctypes mirror of the structs matches the
layout the C compiler gives them.  No compute calls (no GPU here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import smectic3d_amd as S
from smectic3d_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "smectic.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(smx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert "smx_create" in names and "smx_step" in names and "smx_surftrack" in names
    assert sorted(S.EXPORTS) == names


def test_library_exports_every_declared_symbol():
    lib = S.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", S._lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (smx_[a-z0-9_]+)", out))
    assert set(declared_functions()) <= exported


def test_struct_layout_matches_c(tmp_path):
    fields = [f for f, _ in A.SmxConfig._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){",
           'printf("%zu\\n", sizeof(smx_config));']
    src += [f'printf("%zu\\n", offsetof(smx_config, {f}));' for f in fields]
    src += ['printf("%zu %zu %zu\\n", sizeof(smx_surf_opts), offsetof(smx_surf_opts, arm), '
            'offsetof(smx_surf_opts, vel_sign));', "return 0;}"]
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(c), "-o", str(exe)], check=True)
    vals = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert int(vals[0]) == C.sizeof(A.SmxConfig)
    for f, off in zip(fields, vals[1:1 + len(fields)]):
        assert getattr(A.SmxConfig, f).offset == int(off), f
    so, oa, ov = map(int, vals[-3:])
    assert so == C.sizeof(A.SmxSurfOpts) and oa == A.SmxSurfOpts.arm.offset and ov == A.SmxSurfOpts.vel_sign.offset


def test_abi_constants_match_header():
    txt = open(HEADER).read()
    for name, val in re.findall(r"#define (SMX_[A-Z0-9_]+) \(?(-?\d+)\)?", txt):
        assert getattr(A, name) == int(val), name


def test_config_default_matches_reference_constants():
    """smx_config_default restates periodicFFT/pf3d.cpp:35-75 (no device needed)."""
    cfg = A.SmxConfig()
    S.lib().smx_config_default(C.byref(cfg), 256)
    import math
    assert (cfg.nx, cfg.ny, cfg.nz) == (256, 256, 256)
    assert cfg.eps == -0.7 and cfg.beta == 2.0 and cfg.alpha == 1.0 and cfg.gamma == 1.0 and cfg.q0 == 1.0
    assert cfg.dt == 0.001 and cfg.eta == 1.0 and cfg.rho == 1.0 and cfg.advect == 1
    assert cfg.lx == pytest.approx(256 * 2 * math.pi / 16)


def test_invalid_config_fails_before_touching_the_device():
    with pytest.raises(S.SmxError) as e:
        S.Solver(A.make_config(20))
    assert e.value.code == A.SMX_ERR_UNSUPPORTED
    bad = A.make_config(16)
    bad.abi_version = 99
    with pytest.raises(S.SmxError) as e:
        S.Solver(bad)
    assert e.value.code == A.SMX_ERR_INVALID
