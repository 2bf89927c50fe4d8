"""Builds libsbp.so (gfx950) in-tree from sc_2022_amd/csrc.

The stencil pass is instantiated once per tile shape in its own translation
unit (csrc/sbp_tile.hip with -DSBP_TI/-DSBP_TJ); those and the API unit are
compiled in parallel and linked into one shared library with a plain C ABI
(include/sbp.h).

    python -m sc_2022_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libsbp.so"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SBP_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 ops are kept off (plain scalar FP32 arithmetic only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def tile_shapes(name="SBP_TILE_SHAPES"):
    """(ti, tj, aligned) of every compiled stencil tile shape."""
    text = (CSRC / "sbp_launch.h").read_text()
    line = re.search(r"#define " + name + r"\(X\)(.*)", text).group(1)
    return [(int(a), int(b), int(c)) for a, b, c in re.findall(r"X\((\d+),\s*(\d+),\s*(\d+)\)", line)]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{err}")
    return err


def _up_to_date(obj: Path) -> bool:
    """Whether obj is newer than every file its dependency list names."""
    dep = obj.with_suffix(".d")
    if not obj.exists() or not dep.exists():
        return False
    files = dep.read_text().replace("\\\n", " ").split(":", 1)[1].split()
    t = obj.stat().st_mtime
    return all(Path(f).exists() and Path(f).stat().st_mtime <= t for f in files)


def build(jobs: int | None = None, verbose: bool = False, incremental: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    units = [(CSRC / "sbp_api.hip", OBJ / "sbp_api.o", []),
             (CSRC / "sbp_fused.hip", OBJ / "sbp_fused.o", []),
             (CSRC / "sbp_var_fused.hip", OBJ / "sbp_var_fused.o", [])]
    hi = set(tile_shapes("SBP_HI_TILE_SHAPES"))
    for ti, tj, ta in tile_shapes():
        units.append((CSRC / "sbp_tile.hip", OBJ / f"sbp_tile_{ti}x{tj}{'a' if ta else ''}.o",
                      [f"-DSBP_TI={ti}", f"-DSBP_TJ={tj}", f"-DSBP_TA={ta}", f"-DSBP_HI={int((ti, tj, ta) in hi)}"]))
    jobs = jobs or min(len(units), os.cpu_count() or 4)
    cmds = [[HIPCC, *FLAGS, *defs, "-MMD", "-MF", str(obj.with_suffix(".d")), "-c", str(src), "-o", str(obj)]
            for src, obj, defs in units if not (incremental and _up_to_date(obj))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for _, o, _ in units], "-L/opt/rocm/lib", "-lrocsparse", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only the units whose sources or headers changed since their last build")
    a = ap.parse_args(argv)
    print(build(a.jobs, a.verbose, a.incremental))


if __name__ == "__main__":
    main(sys.argv[1:])
