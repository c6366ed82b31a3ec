"""Host-only checks of the chirp-z ring kernels (`csrc/ssg_ring_fft.hip`) by CPU emulation.

The kernel source is compiled for the host against `tests/emu/hip/hip_runtime.h`: one
std::thread per GPU thread, std::barrier for the block and wave ordering points, a
per-wave exchange for `__shfl_xor` and the MFMA products, and the launch's dynamic
LDS size enforced (a write past it aborts).  The build uses AddressSanitizer, so an
out-of-bounds global access by any block, padding blocks included, fails the run.

- `emu_syn`: k_syn_fft against a direct HC2R sum with the azimuthal derivative
  families (long double), per column, in FP64 and in the FP32 form's FP32
  arithmetic (2e-6 per column), and k_syn_cheb against a direct contraction
  (`src/Fourier.jl:463-468, 495-503`, `src/Chebyshev.jl:473-535, 636-722`);
- `emu_ana`: k_ana_fft against a direct R2HC / N with the forward phase filter
  (`src/Fourier.jl:388-395`, `src/phase_filter.jl:105-143`);
- `emu_batch`: the batched launches over RL, RLZ, SL and SLZ grids with job ranges
  padded to multiples of 8 blocks, bitwise against the single-grid launches (FP64
  and FP32 synthesis, both Chebyshev stages, the analysis on double and float input).

Rings with patchOffsetL = 600 reach the 4096-point transforms (three radix-16 passes,
block-wide ordering); the small grids cover the smaller size classes, whose transforms
share a block and order their exchanges per wave.
"""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "springsteel.jl_amd", "csrc")
EMU = os.path.join(REPO, "tests", "emu")
CLANG = "/opt/rocm/llvm/bin/clang++"
FLAGS = ["-std=c++20", "-O1", "-g", "-pthread", "-w", "-fsanitize=address", "-fno-omit-frame-pointer",
         "-D_GLIBCXX_SANITIZE_VECTOR", "-I" + EMU, "-I" + CSRC, "-I" + os.path.join(REPO, "include")]


def _cc(args, cwd):
    r = subprocess.run([CLANG] + FLAGS + args, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    if not os.path.exists(CLANG):
        pytest.skip("no host clang++ for the emulation build")
    d = str(tmp_path_factory.mktemp("emu"))
    # the kernel source as a host translation unit; its quoted includes resolve to the
    # emulation's ssg_mfma.hpp first (the copy lives outside csrc), and its register
    # constraint for the compiler-only value barrier becomes a host register class
    src = open(os.path.join(CSRC, "ssg_ring_fft.hip")).read().replace('"+v"', '"+r"')
    open(os.path.join(d, "ring_fft.cpp"), "w").write(src)
    objs = []
    for name, path in (("ring_fft", os.path.join(d, "ring_fft.cpp")), ("setup", os.path.join(CSRC, "ssg_setup.cpp")),
                       ("rt", os.path.join(EMU, "emu_rt.cpp"))):
        _cc(["-c", path, "-o", name + ".o"], d)
        objs.append(name + ".o")
    for drv in ("emu_syn", "emu_ana", "emu_batch"):
        _cc([os.path.join(EMU, drv + ".cpp")] + objs + ["-o", drv], d)
    return d


def _run(emu, *args, timeout=600):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")
    r = subprocess.run([os.path.join(emu, args[0])] + [str(a) for a in args[1:]], capture_output=True, text=True,
                       timeout=timeout, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    return r.stdout


@pytest.mark.parametrize("geom,nc,nvars,off,spread", [
    ("RL", 10, 3, 0, 0), ("RLZ", 12, 2, 0, 1), ("SL", 12, 2, 0, 0), ("SLZ", 10, 1, 0, 1), ("RL", 4, 1, 600, 1),
])
def test_emulated_synthesis(emu, geom, nc, nvars, off, spread):
    out = _run(emu, "emu_syn", geom, nc, nvars, off, spread)
    assert "worst column relL2" in out


@pytest.mark.parametrize("geom,nc,nvars,off,spread", [("RLZ", 12, 2, 0, 1), ("SL", 12, 2, 0, 0), ("RL", 4, 1, 600, 1)])
def test_emulated_synthesis_fp32(emu, geom, nc, nvars, off, spread):
    # the FP32 form's transforms: float input, tables and output, FP32 arithmetic
    out = _run(emu, "emu_syn", geom, nc, nvars, off, spread, 1)
    assert "worst column relL2" in out


@pytest.mark.parametrize("geom,nc,nvars,off,cap", [
    ("RL", 10, 3, 0, -1), ("RLZ", 12, 2, 0, -1), ("SLZ", 10, 2, 0, 3), ("SL", 12, 1, 0, -1), ("RL", 4, 1, 600, -1),
])
def test_emulated_analysis(emu, geom, nc, nvars, off, cap):
    out = _run(emu, "emu_ana", geom, nc, nvars, off, cap)
    assert "worst column relL2" in out


def test_emulated_batch_padding_blocks(emu):
    out = _run(emu, "emu_batch")
    assert out.count("bitwise") == 20 and "DIFFERS" not in out
