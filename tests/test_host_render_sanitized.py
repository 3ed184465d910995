"""The library's host code for the render, on the CPU under AddressSanitizer
and UBSan: the .hip sources compiled host-only and linked against a CPU
stand-in of the HIP runtime (tests/cpp/hip_runtime_stub.cpp) in place of
libamdhip64, driven by tests/cpp/render_host_driver.cpp through the C ABI.

Kernels do not run here (the GPU tests check outputs); what runs is the
planner, the table uploads, every launch sequence, the chunk groups' forks
and joins (staggered and side by side), stream capture, graph
instantiation, upload and replay, profiling, the surrogate form and the
autotuner, on the gigapixel canvas (50 chunks) and the 512 uneven chunks.
The stub checks each runtime call (copies and fills inside live
allocations, captures joined before they end, no synchronising or
allocating call and no memset node during capture, live graphs and
events); a separate program shows that each of those checks fires."""
import concurrent.futures
import os
import shutil
import subprocess

import numpy as np
import pytest

from image360_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "image360_amd", "csrc")
CPP = os.path.join(ROOT, "tests", "cpp")
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"
SOURCES = ["render_float", "capi", "color", "export", "surrogate", "exact_ops", "blend_exact", "profile"]
# host-only compilation: no device code is generated, and the sanitizers
# instrument the host side only
FLAGS = ["-O1", "-g", "-std=c++17", "-fPIC", "--offload-host-only", "-x", "hip", "-w", "-fno-omit-frame-pointer",
         "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
         "-Xarch_host", "-fno-sanitize-recover=all", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)), reason="ROCm toolchain absent")


def _compile(src, obj):
    r = subprocess.run([HIPCC, *FLAGS, "-c", src, "-o", obj], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return obj


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("host_render")
    jobs = [(os.path.join(CSRC, f"{s}.hip"), str(d / f"{s}.o")) for s in SOURCES]
    jobs += [(os.path.join(CPP, f"{s}.cpp"), str(d / f"{s}.o"))
             for s in ("hip_runtime_stub", "render_host_driver", "stub_teeth")]
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))
    # each host-only object refers to its (absent) device code object by a
    # per-file symbol; define them empty
    r = subprocess.run(["nm", "-u", *objs], capture_output=True, text=True, check=True)
    fat = sorted({w for w in r.stdout.split() if w.startswith("__hip_fatbin_")})
    assert fat, "no device code object symbols: the sources were not compiled as HIP"
    (d / "fatbins.cpp").write_text("".join(f'extern "C" const unsigned char {s}[16] = {{0}};\n' for s in fat))
    subprocess.run([CLANG, "-c", "-fPIC", str(d / "fatbins.cpp"), "-o", str(d / "fatbins.o")], check=True)
    lib_objs = objs[:len(SOURCES)] + [str(d / "fatbins.o")]
    stub = str(d / "hip_runtime_stub.o")
    exes = {}
    for name, main in (("driver", str(d / "render_host_driver.o")), ("teeth", str(d / "stub_teeth.o"))):
        exe = str(d / name)
        link = [CLANG, "-fsanitize=address,undefined", "-o", exe, main, stub]
        link += lib_objs if name == "driver" else []
        r = subprocess.run(link + ["-lpthread"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        exes[name] = exe
    yield d, exes
    shutil.rmtree(d, ignore_errors=True)


def _geometry(g, path, bands):
    with open(path, "w") as f:
        f.write(f"{g.num_frames} {len(g.chunks)} {bands}\n")
        for i in range(g.num_frames):
            h = " ".join(f"{v:.17g}" for v in np.asarray(g.H[i], np.float64).reshape(-1))
            f.write(f"{int(g.frame_w[i])} {int(g.frame_h[i])} {float(g.gains[i]):.9g} {h}\n")
        for c in g.chunks:
            f.write(f"{c.x} {c.y} {c.w} {c.h} {len(c.contributors)} " + " ".join(str(int(v)) for v in c.contributors)
                    + "\n")


def _run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("I360HIP_RENDER_GROUPS", None)
    env.pop("I360HIP_CHAMFER_STRIP", None)
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=600, env=env)


@pytest.mark.parametrize("workload", ["gigapixel", "uneven"])
def test_render_host_code_under_sanitizers(harness, workload):
    d, exes = harness
    g = synth.gigapixel()
    bands = 7
    if workload == "uneven":
        g, bands = synth.uneven_chunks(base=g), 6
    path = str(d / f"{workload}.txt")
    _geometry(g, path, bands)
    r = _run(exes["driver"], path)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-6000:]
    assert "driver done" in r.stdout and "HIP STUB OK" in r.stdout
    assert "runtime error" not in r.stderr
    # the graph paths ran: captures and replays, every chunk group split
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("HIP STUB OK")][0]
    n = [int(w) for w in line.replace(",", " ").split() if w.isdigit()]
    launches, captures, graph_launches = n[0], n[1], n[2]
    assert launches > 1000 and captures >= 8 and graph_launches >= 8, line


@pytest.mark.parametrize("mistake,message", [
    ("unjoined", "is not joined into the origin"),
    ("memset_in_capture", "hipMemsetAsync recorded into a captured graph"),
    ("copy_overrun", "runs 1 bytes past its allocation"),
    ("free_inside", "is not the start of a live allocation"),
    ("sync_in_capture", "called while a stream capture is active"),
])
def test_runtime_stub_catches_mistakes(harness, mistake, message):
    _, exes = harness
    ok = _run(exes["teeth"], "ok")
    assert ok.returncode == 0 and "no violation" in ok.stdout, ok.stderr[-2000:]
    r = _run(exes["teeth"], mistake)
    assert r.returncode != 0 and message in r.stderr, (r.returncode, r.stderr[-2000:])
