"""Tile-shape autotuner for the stencil pass.

For every multigrid level side N_l of a problem, every SBP order and both
precisions, times the matrix-free apply at each compiled tile shape (TI
threads along i x J rows per thread) with HIP events on the launch stream,
over input/output pairs rotated so that the working set exceeds the 256 MB
Infinity Cache when the level is that large, and keeps the fastest shape.

Results are stored in tiles_gfx950.json next to this file, keyed
"p{p}/{f64|f32}/{N}"; SbpPoisson applies them per level at construction.

    python -m sc_2022_amd.autotune --n 4097 --p 2 4
"""
from __future__ import annotations

import argparse
import json
import os
import re
from pathlib import Path

HERE = Path(__file__).resolve().parent
CACHE = HERE / "tiles_gfx950.json"


def compiled_shapes(p=2):
    """Tile shapes (ti, tj, aligned) libsbp.so has for order p (p = 6, 8, 10: a subset)."""
    text = (HERE / "csrc" / "sbp_launch.h").read_text()
    name = "SBP_TILE_SHAPES" if p <= 4 else "SBP_HI_TILE_SHAPES"
    line = re.search(r"#define " + name + r"\(X\)(.*)", text).group(1)
    return [(int(a), int(b), int(c)) for a, b, c in re.findall(r"X\((\d+),\s*(\d+),\s*(\d+)\)", line)]


def shape_name(shape):
    """'256x8' for an unaligned shape, '256x8a' for its line-aligned variant."""
    return f"{shape[0]}x{shape[1]}" + ("a" if len(shape) > 2 and shape[2] else "")


def load_cache(path: os.PathLike | None = None) -> dict:
    p = Path(path) if path else CACHE
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def key(p, dtype_tag, N):
    return f"p{p}/{dtype_tag}/{N}"


def lookup(p, dtype_tag, N, cache=None):
    c = load_cache() if cache is None else cache
    v = c.get(key(p, dtype_tag, N))
    return tuple(v["tile"]) if v else None


def time_level(h, level, dtype, shapes, reps=None):
    import torch
    N = h.level_size(level)
    n = N * N
    es = 8 if dtype == torch.float64 else 4
    nbuf = max(2, min(16, int((1 << 30) // max(1, 2 * n * es)) + 1))
    reps = reps or max(20, min(400, int(2e9 // max(1, 2 * n * es))))
    ins = [torch.randn(n, dtype=dtype, device=h.device) for _ in range(nbuf)]
    outs = [torch.empty_like(ins[0]) for _ in range(nbuf)]
    res = {}
    for shape in shapes:
        ti, tj = shape[0], shape[1]
        h.set_tiles(level, ti, tj, aligned=len(shape) > 2 and bool(shape[2]))
        for k in range(nbuf):
            h.apply(ins[k], outs[k], level=level)
        torch.cuda.synchronize()
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for r in range(reps):
            h.apply(ins[r % nbuf], outs[r % nbuf], level=level)
        e1.record(s)
        torch.cuda.synchronize()
        res[tuple(shape)] = e0.elapsed_time(e1) / reps * 1e3   # microseconds
    return res


def tune(N, orders=(2, 4), out=None, verbose=True):
    import torch
    from .solver import SbpPoisson
    cache = load_cache(out)
    for p in orders:
        shapes = compiled_shapes(p)
        h = SbpPoisson(N, p)
        nlev = max(1, h.nlevels)
        for level in range(nlev):
            Nl = h.level_size(level)
            # line-aligned shapes only where a level streams far more than launch latency
            lshapes = [sh for sh in shapes if Nl >= 2049 or not sh[2]]
            for dtype, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
                res = time_level(h, level, dtype, lshapes)
                best = min(res, key=res.get)
                es = 8 if tag == "f64" else 4
                cache[key(p, tag, Nl)] = {"tile": list(best), "us": res[best],
                                          "GBs": 2 * Nl * Nl * es / (res[best] * 1e-6) / 1e9,
                                          "all_us": {shape_name(k): round(v, 3) for k, v in res.items()}}
                if verbose:
                    print(f"p={p} {tag} N={Nl}: best {best} {res[best]:.2f} us "
                          f"({cache[key(p, tag, Nl)]['GBs']:.0f} GB/s)", flush=True)
        h.close()
    path = Path(out) if out else CACHE
    path.write_text(json.dumps(cache, indent=1, sort_keys=True))
    return cache


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, nargs="+", default=[4097])
    ap.add_argument("--p", type=int, nargs="+", default=[2, 4])
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    for N in a.n:
        tune(N, tuple(a.p), a.out)


if __name__ == "__main__":
    main()
