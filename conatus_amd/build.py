"""Build recipe of libcnhip.so (hand-written gfx950 HIP kernels behind a C ABI).

``python -m conatus_amd.build`` compiles every ``csrc/*.hip`` into ``conatus_amd/libcnhip.so`` with
hipcc for gfx950 only.  FP contraction is off (the kernels place their fma explicitly) and the
packed-FP32 target feature is disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
LIB = PKG / "libcnhip.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or Path(cand).exists()):
            return cand
    raise RuntimeError("hipcc not found")


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


#: banded_lane.hip is compiled once per (element type, block length, border rows, kl): one object
#: each, so that the lane-mode kernels build in parallel
LANE_TYPES = (("zc", 1), ("d", 0))
LANE_SPX = (16, 8)
LANE_RB = (2, 8)
LANE_KL = (1, 2, 4, 6)


def compile_units() -> list[tuple[Path, str, list[str]]]:
    """(source, object stem, extra flags) of every object of the library."""
    units = []
    for src in sources():
        if src.name != "banded_lane.hip":
            units.append((src, src.stem, []))
            continue
        for tag, zc in LANE_TYPES:
            for spx in LANE_SPX:
                for rb in LANE_RB:
                    for kl in LANE_KL:
                        t = f"{tag}{spx}_r{rb}"
                        flags = [f"-DCN_LANE_ZC={zc}", f"-DCN_LANE_SPX={spx}", f"-DCN_LANE_RB={rb}",
                                 f"-DCN_LANE_KL={kl}", f"-DCN_LANE_SOLVE=lane_solve_{t}_k{kl}"]
                        if kl == LANE_KL[0]:
                            flags.append(f"-DCN_LANE_BORDER=lane_border_{t}")
                        units.append((src, f"banded_lane_{t}_k{kl}", flags))
    return units


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd[-3:])}")


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile the library (always: a stale .so must not survive a source change).  Every source is
    compiled to an object in parallel (jobs, default: the usable cores), then linked."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    units = compile_units()
    objs = [objdir / (stem + ".o") for _, stem, _ in units]
    flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    cmds = [[hipcc(), *flags, *extra, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
            for (src, _, extra), obj in zip(units, objs)]
    njobs = jobs or max(1, min(len(cmds), len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(max_workers=njobs) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    p = build(verbose=True)
    print(f"built {p}")
