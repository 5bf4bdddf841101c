"""The compiled wave-streamed SOR kernel (core_collapse_amd/csrc/sorws.hip, gfx950 assembly from hipcc),
executed on the CPU by tests/isa/interp.py with several interleaved waves.

The kernel source is also modelled in test_sor_schedule.py; this runs the machine code itself, so that a
transformation of the compiler (for example a structurized loop that sends some lanes back to the loop
header without the others) shows up on the CPU.  Checked: every run ends within its instruction budget,
raises no wait timeout, touches no address outside its buffers, and gives the lexicographic sweep (FP64,
to FMA rounding) or the block-refined mixed sweep of test_sor_schedule.py (to FP32 rounding).  Grids cover
edge tiles, interior tiles that take the select-free path, partial sweep blocks and one-row grids.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from tests.isa.interp import Machine, parse_kernel, run_waves
from tests.test_sor_schedule import block_refined_sor, lex_sor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "--cuda-device-only", "-S"]


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("isa") / "sorws.s")
    subprocess.check_call([HIPCC] + FLAGS + [os.path.join(ROOT, "core_collapse_amd", "csrc", "sorws.hip"), "-o", out],
                          stderr=subprocess.DEVNULL)
    return out


def _run(asm, Nx, Ny, nsw, SW, mixed, nwaves, seed=0):
    name = "_ZN2cc8k_sor_wsI%sLi%dEEEvNS_6WsArgsE" % ("f" if mixed else "d", SW)
    prog, labels = parse_kernel(asm, name)
    NB = (Ny + SW - 1 + 63) // 64
    pitch = 64 * NB + 2
    K = Nx + 64 * NB + SW + 16
    ntiles = NB * ((nsw + SW - 1) // SW)
    RN = 2 * SW + (4 if mixed else 0)
    rec_stride = (Nx + SW + 160) * RN
    rsz = 4 if mixed else 8
    m = Machine(64 << 20)
    D = m.alloc(K * pitch * 8)
    C = m.alloc(K * pitch * 8)
    rec = m.alloc(ntiles * rec_stride * rsz)
    state = m.alloc((ntiles + 2) * 4)
    karg = m.alloc(104)
    rng = np.random.default_rng(seed)
    src = rng.standard_normal((Ny, Nx))
    Cd = np.zeros((K, pitch))
    for r in range(Ny):
        Cd[r:r + Nx, r] = src[r]
    m.mem[C - m.base:C - m.base + K * pitch * 8] = Cd.reshape(-1).view(np.uint8)
    kx, ky, om = 0.3, 0.2, 1.8
    kb = struct.pack("<6i3d4QI4x2Q", Nx, Ny, nsw, NB, pitch, ntiles, kx, ky, om, D, C, rec, rec_stride,
                     rec_stride * rsz, state, 10_000_000)
    m.mem[karg - m.base:karg - m.base + len(kb)] = np.frombuffer(kb, np.uint8)
    ok, waves, total = run_waves(m, prog, labels, karg, nwaves, 20_000_000, rng)
    assert ok, ("instruction budget exceeded", [(w.done, w.pc) for w in waves])
    assert m.rd32(state + 4) == 0, "a wait timed out"
    Dd = m.mem[D - m.base:D - m.base + K * pitch * 8].view(np.float64).reshape(K, pitch)
    x = np.array([Dd[r:r + Nx, r] for r in range(Ny)])
    want = block_refined_sor(src, nsw, SW, kx, ky, om) if mixed else lex_sor(src, nsw, kx, ky, om)
    return np.abs(x - want).max() / np.abs(want).max()


@pytest.mark.parametrize("Nx,Ny,nsw,SW,mixed,nwaves", [(20, 9, 9, 4, False, 3), (90, 200, 8, 4, False, 4),
                                                        (33, 65, 17, 8, True, 3), (90, 200, 9, 4, True, 4),
                                                        (97, 2, 9, 4, False, 2), (2, 97, 9, 8, True, 3)])
def test_compiled_kernel_runs_the_sweep(asm, Nx, Ny, nsw, SW, mixed, nwaves):
    err = _run(asm, Nx, Ny, nsw, SW, mixed, nwaves)
    assert err < (1e-6 if mixed else 1e-14), err
