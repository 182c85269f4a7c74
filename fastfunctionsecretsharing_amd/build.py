"""Builds the device library ``libfastfss_hip.so`` (gfx950) in-tree.

Every ``csrc/*.hip`` and ``csrc/*.cpp`` is compiled with hipcc for
``--offload-arch=gfx950`` in parallel, then linked into one shared library
exporting the ``FastFss_*`` C ABI declared under ``include/FastFss``.
``probe/lds_probe.hip`` becomes ``libfss_lds_probe.so``, the LDS read-rate
probe the benchmark measures its roofline ceiling with (not part of the C ABI).

Usage: ``python -m fastfunctionsecretsharing_amd.build [-j N]``.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libfastfss_hip.so"
PROBE_SRC = PKG / "probe" / "lds_probe.hip"
PROBE_OUT = PKG / "libfss_lds_probe.so"
BUILD = PKG / "_build"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FSS_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-DFAST_FSS_BUILD",
    "-fvisibility=hidden",
    f"-I{REPO / 'include'}",
    f"-I{CSRC}",
    # packed FP32 VALU forms are disabled for all device code of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# Per-source flags: LLVM's iterative ILP scheduler, where same-process A/B runs
# showed it faster with equal outputs.
#  - spline_eval.hip: a wave keeps up to 16 AES table lookups in flight in the
#    spline walk instead of 1-3 (7% faster at 2^22 elements, 35% at 1024).  The
#    option crashes the compile of spline.hip's keyGen kernels (and of dcf.hip
#    and mic.hip), which is why the eval launchers have a file of their own.
#  - dpf_grotto.hip: DPF eval of 2^14 elements 10% faster, 2^24 and evalAll
#    0.3-1%, full-domain Grotto unchanged.
#  Measured and left on the default scheduler: ottt_prng.hip (OTTT 16-26%
#  slower) and batch.hip (ragged batch unchanged).
ILP_SCHEDULER = ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"]
SOURCE_FLAGS = {
    "spline_eval.hip": ILP_SCHEDULER,
    "dpf_grotto.hip": ILP_SCHEDULER,
}


def sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    flags = [*CFLAGS, *SOURCE_FLAGS.get(src.name, [])]
    cmd = [HIPCC, *flags, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC, *flags, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"hipcc failed on {src.name}:\n{msg}")
    return obj


def _up_to_date(src: Path) -> bool:
    obj = BUILD / (src.name + ".o")
    if not obj.exists():
        return False
    deps = [src, *CSRC.glob("*.h"), *(REPO / "include").rglob("*.h"), Path(__file__)]
    return obj.stat().st_mtime >= max(d.stat().st_mtime for d in deps)


def build(jobs: int | None = None, verbose: bool = True, incremental: bool = False) -> Path:
    """Compile every source (or, with ``incremental``, those whose object is older
    than the source, any header or this file) and link the library."""
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    todo = [s for s in srcs if not (incremental and _up_to_date(s))]
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [BUILD / (s.name + ".o") for s in srcs]
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if not (incremental and PROBE_OUT.exists() and PROBE_OUT.stat().st_mtime >= PROBE_SRC.stat().st_mtime):
        ptmp = PROBE_OUT.with_suffix(".so.tmp")
        cmd = [HIPCC, *CFLAGS, "-shared", str(PROBE_SRC), "-o", str(ptmp)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
            raise RuntimeError(f"hipcc failed on {PROBE_SRC.name}:\n{msg}")
        os.replace(ptmp, PROBE_OUT)
    if verbose:
        print(f"built {OUT.relative_to(REPO)} from {len(srcs)} sources and {PROBE_OUT.relative_to(REPO)}")
    return OUT


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only sources whose object is older than the source or any header")
    a = ap.parse_args(argv)
    try:
        build(a.j, incremental=a.incremental)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
