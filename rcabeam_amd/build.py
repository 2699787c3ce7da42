"""Build the MI355X shared library `rcabeam_amd/lib/librcabeam_mi355x.so` with hipcc.

Usage: python -m rcabeam_amd.build [-v] [--incremental]  (or `__graft_entry__.build()`).
`--incremental` recompiles only sources newer than their object (or than any header).

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
generation disabled (`-mno-packed-fp32` equivalent feature flag), and linked
with rocSOLVER/rocBLAS for the Hermitian eigendecomposition.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
OUT_DIR = PACKAGE / "lib"
LIBRARY = OUT_DIR / "librcabeam_mi355x.so"
ARCH = "gfx950"

SOURCES = ["pack.hip", "das.hip", "das_tile.hip", "das_mfma.hip", "das_grid.hip", "lowrank.hip", "ingest.hip", "two_stage.hip", "combine.hip", "hybrid.hip", "frame.hip", "clutter.hip", "svd.hip", "demod.hip", "scene.hip"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-ffp-contract=fast",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    candidate = Path(rocm) / "bin" / "hipcc"
    return str(candidate) if candidate.exists() else (shutil.which("hipcc") or "hipcc")


def build(verbose: bool = False, incremental: bool = False) -> Path:
    """Compile all sources and link the shared library; return its path.

    With `incremental`, a source whose object is newer than the source and every header is not
    recompiled (the default rebuilds everything)."""
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    obj_dir = ROOT / "_build"
    obj_dir.mkdir(exist_ok=True)
    hipcc = _hipcc()
    includes = [f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    sources = [CSRC / s for s in SOURCES if (CSRC / s).exists()]

    headers = [*CSRC.glob("*.hpp"), *(ROOT / "include").glob("*.h")]
    newest_header = max((h.stat().st_mtime for h in headers), default=0.0)

    def compile_one(src: Path) -> Path:
        obj = obj_dir / (src.stem + ".o")
        if incremental and obj.exists() and obj.stat().st_mtime > max(src.stat().st_mtime, newest_header):
            return obj
        cmd = [hipcc, *HIPCC_FLAGS, *includes, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        result = subprocess.run(cmd, capture_output=True, text=True)
        if result.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{result.stderr}")
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        objects = list(pool.map(compile_one, sources))
    tmp = LIBRARY.with_suffix(".so.tmp")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = [
        hipcc,
        f"--offload-arch={ARCH}",
        "-shared",
        "-fPIC",
        *map(str, objects),
        f"-L{rocm}/lib",
        "-lrocsolver",
        "-lrocblas",
        f"-Wl,-rpath,{rocm}/lib",
        "-o",
        str(tmp),
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed:\n{result.stderr}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv)
    print(path)
