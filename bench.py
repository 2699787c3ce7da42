"""Benchmark: filtered OPW power Doppler (OPW + SVD clutter filter + PD) on one MI355X.

Workload: 128 + 128-element RCA at 200 um, c = 1540 m/s, fs = 20 MHz,
f_demod = 15 MHz, RC and CR complex64 channel IQ (1024 samples, 128 channels,
11 angles uniform in +/-10 deg, 200 frames), grid 128 x 256 x 128 over
+/-12.8 mm x 4-20 mm, F = 1, clutter rank_start = 3.  Scene (generated on
the GPU, seeded): 20 000 static Rayleigh tissue scatterers with a slow common
phase drift, 2 000 blood scatterers at -40 dB moving at 1-20 mm/s in three
tube segments, -60 dB noise.

One step = pack RC and CR -> frame-batched OPW delay-and-sum (complex64 IQ in
HBM) -> scaled Gram on MFMA -> float64 eigendecomposition -> filtered power,
starting from channel data resident in HBM.  `--precision float16` (or
`bfloat16`) rounds channels and beamformed IQ to 16 bits, delay-and-sum on
MFMA with one 16-bit coefficient per sample (exact products, float32
accumulation), Gram formed exactly from the 16-bit IQ.  `float16-lowrank`
(or `bfloat16-lowrank`) is the mixed-precision form reported beside the
float32 one: a channel Gram (split bfloat16 hi + lo products on MFMA over every
4th channel row) and a float64 eigendecomposition give the 4 strongest
temporal modes of the channels, which are coded as 16-bit hi + lo
coefficients while only the residual is rounded to 16 bits (rcabeam_amd.lowrank),
MFMA delay-and-sum with 16-bit coefficients, complex64 IQ, IQ Gram with split
bfloat16 products on MFMA.

Prints one JSON line.  `value` is voxel-frames/s of the whole step.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0  # dense bf16 / fp16 MFMA
FLOPS_PER_TAP = 14  # 2 sub + 2 fma (interpolation) + 4 fma (rotate-accumulate), as counted in the reference kernel
PROFILES = ROOT / "profiles"
# The delay-and-sum kernel each precision's rcabeam_opw_grid_device launch runs on this workload.
DAS_KERNEL = {
    "float32": "das_opw_grid_valu<13, false, false, true> (complex64 channels, float32 VALU)",
    "float16": "das_opw_mfma<__half2> (float16 channels, MFMA)",
    "bfloat16": "das_opw_mfma<bf16x2> (bfloat16 channels, MFMA)",
    "float16-lowrank": "das_opw_mfma<__half2> (low-rank coded float16 channels, MFMA)",
    "bfloat16-lowrank": "das_opw_mfma<bf16x2> (low-rank coded bfloat16 channels, MFMA)",
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--precision", choices=("float32", "float16", "bfloat16", "float16-lowrank", "bfloat16-lowrank"),
                   default="float32",
                   help="float32: complex64 channels and IQ; float16 / bfloat16: 16-bit channels and IQ; "
                        "*-lowrank: low-rank + 16-bit residual channels, complex64 IQ (the mixed-precision form)")
    p.add_argument("--lowrank-modes", type=int, default=4, help="temporal modes k of the low-rank channel coding")
    p.add_argument("--grid", type=int, nargs=3, default=(128, 256, 128))
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision comparison runs")
    p.add_argument("--no-combiners", action="store_true", help="skip the full-size combiner timings")
    p.add_argument("--no-two-stage", action="store_true", help="skip the full-size two-stage approximation timings")
    p.add_argument("--two-stage-only", action="store_true",
                   help="only time the two-stage approximation on the full-size scene (profiling runs)")
    p.add_argument("--combiners-only", action="store_true",
                   help="only time the combiners on the full-size scene (profiling runs)")
    p.add_argument("--aperture-passes-only", action="store_true",
                   help="only time the float32 delay-and-sum as its two per-aperture passes (profiling runs)")
    p.add_argument("--cpu-voxels", type=int, default=600,
                   help="full-size voxels at which each form is compared with the float64 oracle (0: skip)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a plain 1 GiB device-to-device copy three times (HBM counter calibration)")
    return p.parse_args()


def copy_calibration() -> int:
    """Three plain copies of 1 GiB: the reference point for FETCH_SIZE / WRITE_SIZE scaling."""
    import torch

    n = 1 << 30
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "copies": 3}), flush=True)
    return 0


class Pipeline:
    """Preallocated buffers and the one-step enqueue of the filtered-OPW pipeline."""

    def __init__(self, scene, rc, cr, storage: str, rank_start: int, lowrank_modes: int = 4):
        import torch

        from rcabeam_amd import _device as dev
        from rcabeam_amd import _lib
        from rcabeam_amd.ensemble import _Geometry

        self.torch = torch
        self.dev = dev
        self.lib = _lib.load()
        self.rc, self.cr = rc, cr
        self.lowrank = storage.endswith("-lowrank")
        self.channel_storage = storage.replace("-lowrank", "")
        self.storage = storage
        # beamformed IQ storage: complex64 for float32 and the low-rank forms
        iq_storage = "float32" if (self.lowrank or storage == "float32") else storage
        self.code = _lib.STORAGE[iq_storage]
        self.rank_start = rank_start
        self.g = _Geometry(scene.angles, scene.t_start, scene.grid, scene.geom)
        self.ns, self.nc, self.na, self.nt = rc.shape
        nv = self.g.nv
        self.k = lowrank_modes if self.lowrank else 0
        self.ntc = self.nt + 2 * self.k  # frames of the packed channels
        words = torch.complex64 if storage == "float32" else torch.int32
        self.packed = [dev.empty((self.nc, self.na, self.ns, self.ntc), words) for _ in range(2)]
        self.range_error = dev.empty((1,), torch.int32)
        self.pd = dev.empty((nv, 2), torch.float32)
        self.signal = dev.empty((nv, self.nt), torch.complex64 if self.code == 0 else torch.int32)
        if self.lowrank:
            from rcabeam_amd.lowrank import ROW_STEP, SUBSPACE_ITERATIONS

            self.row_step = ROW_STEP
            self.sub_iterations = SUBSPACE_ITERATIONS
            rows = (self.ns * self.nc * self.na + ROW_STEP - 1) // ROW_STEP
            self.c_rows = rows
            self.c_work_bytes = int(self.lib.rcabeam_gram_workspace_bytes(rows, self.nt))
            self.c_work = dev.empty((self.c_work_bytes // 4,), torch.float32)
            self.c_stats = dev.empty((3,), torch.int32)
            self.c_real = dev.empty((int(self.lib.rcabeam_gram_real_elements(self.nt)),), torch.float64)
            self.c_gram = dev.empty((self.nt, self.nt), torch.complex128)
            self.basis = dev.empty((self.k, self.nt), torch.complex128)
            self.coded = dev.empty((nv, self.ntc), torch.complex64)
        self.stats = dev.empty((3,), torch.int32)
        self.real = dev.empty((int(self.lib.rcabeam_gram_real_elements(self.nt)),), torch.float64)
        # IQ Gram: exact float64 for the float32 form; split bfloat16 (hi + lo) products on MFMA for the
        # low-rank mixed form; one exact 16-bit product for 16-bit IQ.
        self.exact_gram = storage == "float32"
        if self.exact_gram:
            self.work_bytes = int(self.lib.rcabeam_gram_f64_workspace_bytes(nv, self.nt))
            self.work = dev.empty((self.work_bytes // 8,), torch.float64)
        else:
            self.work_bytes = int(self.lib.rcabeam_gram_workspace_bytes(nv, self.nt))
            self.work = dev.empty((self.work_bytes // 4,), torch.float32)
        self.gram = dev.empty((self.nt, self.nt), torch.complex128)
        self.values = dev.empty((self.nt,), torch.float64)
        self.e_work = dev.empty((self.nt,), torch.float64)
        self.info = dev.empty((1,), torch.int32)
        self.vectors = dev.empty((rank_start, self.nt), torch.complex128)
        self.power = dev.empty((nv,), torch.float64)

    def _call(self, name, *args):
        from rcabeam_amd import _lib

        _lib.check(getattr(self.lib, name)(*args), name)

    def step(self, das_events=None, marks=None, hosts=None):
        """Enqueue one full step on the null stream.  Optional `das_events` bracket the DAS launch;
        `marks` (a list) receives (stage name, event) pairs recorded at the end of each stage.  `hosts`
        (float32 form only): pageable RC / CR arrays uploaded inside the step, CR overlapping the RC pass."""
        from rcabeam_amd import _lib

        dev, nv = self.dev, self.g.nv

        def mark(name):
            if marks is not None:
                ev = self.torch.cuda.Event(enable_timing=True)
                ev.record()
                marks.append((name, ev))

        mark("start")
        c, fs, fd, fn = self.g.constants
        shape = (self.ns, self.nc, self.na, self.nt)
        channel_code = _lib.STORAGE[self.channel_storage]
        if self.lowrank:
            # Channel basis (rcabeam_amd.lowrank.channel_basis): split-bfloat16 Gram of every row_step-th row
            # of both apertures' Casorati matrices, float64 subspace iteration for its k strongest modes.
            for i, raw in enumerate((self.rc, self.cr)):
                self._call("rcabeam_iq_absmax_device", dev.ptr(raw), 0, raw.numel(), dev.ptr(self.c_stats), i)
            for i, raw in enumerate((self.rc, self.cr)):
                self._call("rcabeam_temporal_gram_sampled_device", dev.ptr(raw), 0, self.c_rows, self.row_step,
                           self.nt, dev.ptr(self.c_stats), dev.ptr(self.c_real), dev.ptr(self.c_work),
                           self.c_work_bytes, i)
            self._call("rcabeam_gram_hermitian_device", dev.ptr(self.c_real), self.nt, 1.0 / (2 * self.c_rows),
                       dev.ptr(self.c_gram))
            self._call("rcabeam_dominant_subspace_device", dev.ptr(self.c_gram), self.nt, self.k,
                       self.sub_iterations, dev.ptr(self.basis))
            mark("channel_basis")
            for raw, packed in zip((self.rc, self.cr), self.packed):
                self._call("rcabeam_pack_ensemble_lowrank_device", dev.ptr(raw), dev.ptr(packed), channel_code,
                           dev.ptr(self.basis), self.k, dev.ptr(self.range_error), *shape)
        elif hosts is None:
            for raw, packed in zip((self.rc, self.cr), self.packed):
                if self.storage == "float32":
                    self._call("rcabeam_pack_ensemble_device", dev.ptr(raw), dev.ptr(packed), *shape)
                else:
                    pack = ("rcabeam_pack_ensemble_fp16_device" if self.storage == "float16"
                            else "rcabeam_pack_ensemble_bf16_device")
                    self._call(pack, dev.ptr(raw), dev.ptr(packed), dev.ptr(self.range_error), *shape)
        nx, nz, ny = self.g.shape
        if hosts is not None:
            # Host channels: RC uploaded, packed and summed; CR uploaded on the copy stream while that pass runs,
            # then packed and added (rcabeam_opw_grid_aperture_device), complex64 only.
            threads = dev.upload_threads()
            for aperture, (h, raw, packed) in enumerate(zip(hosts, (self.rc, self.cr), self.packed)):
                name = "rcabeam_upload_host" if aperture == 0 else "rcabeam_upload_host_concurrent"
                self._call(name, h.ctypes.data, dev.ptr(raw), h.nbytes, threads)
                self._call("rcabeam_pack_ensemble_device", dev.ptr(raw), dev.ptr(packed), *shape)
                self._call("rcabeam_opw_grid_aperture_device", dev.ptr(self.packed[0]), dev.ptr(self.packed[1]),
                           dev.ptr(self.g.axes[0]), dev.ptr(self.g.axes[1]), dev.ptr(self.g.axes[2]), nx, nz, ny,
                           dev.ptr(self.g.xe), dev.ptr(self.g.ye), dev.ptr(self.g.angles), dev.ptr(self.g.starts),
                           dev.ptr(self.pd), dev.ptr(self.signal), aperture, aperture, self.ns, self.nc, self.na,
                           self.nt, c, fs, fd, fn)
        else:
            mark("pack")
            if das_events is not None:
                das_events[0].record()
            # OPW on the Cartesian grid (8 x 8 lateral tiles).  float32: complex64 channels and IQ, VALU;
            # 16-bit: MFMA with one 16-bit coefficient per sample (exact products, float32 accumulation).
            coef = 0 if self.storage == "float32" else _lib_coef_16bit()
            out = self.coded if self.lowrank else self.signal
            self._call("rcabeam_opw_grid_device", dev.ptr(self.packed[0]), dev.ptr(self.packed[1]), channel_code,
                       coef, dev.ptr(self.g.axes[0]), dev.ptr(self.g.axes[1]), dev.ptr(self.g.axes[2]), nx, nz, ny,
                       dev.ptr(self.g.xe), dev.ptr(self.g.ye), dev.ptr(self.g.angles), dev.ptr(self.g.starts),
                       dev.ptr(self.pd), dev.ptr(out), 0 if self.lowrank else self.code, 0, self.ns, self.nc,
                       self.na, self.ntc, c, fs, fd, fn)
            if das_events is not None:
                das_events[1].record()
        mark("das")
        if self.lowrank:
            self._call("rcabeam_lowrank_expand_device", dev.ptr(self.coded), dev.ptr(self.basis), self.k,
                       dev.ptr(self.signal), 0, dev.ptr(self.pd), nv, self.nt)
            mark("expand")
        if self.exact_gram:
            # Float64 products and sums of complex64 samples cannot overflow, so the exact Gram needs no
            # |max| scaling pass (its eigenvectors, all the filter uses, do not depend on the scale).
            self._call("rcabeam_temporal_gram_f64_device", dev.ptr(self.signal), nv, self.nt, dev.ptr(self.real),
                       dev.ptr(self.work), self.work_bytes, 0)
        else:
            self._call("rcabeam_iq_absmax_device", dev.ptr(self.signal), self.code, nv * self.nt,
                       dev.ptr(self.stats), 0)
            self._call("rcabeam_temporal_gram_device", dev.ptr(self.signal), self.code, nv, self.nt,
                       dev.ptr(self.stats), dev.ptr(self.real), dev.ptr(self.work), self.work_bytes, 0)
        mark("gram")
        self._call("rcabeam_gram_hermitian_device", dev.ptr(self.real), self.nt, 1.0 / nv, dev.ptr(self.gram))
        self._call("rcabeam_eigh_device", dev.ptr(self.gram), dev.ptr(self.values), dev.ptr(self.e_work),
                   dev.ptr(self.info), self.nt)
        # Discarded vectors = eigenvectors of the rank_start largest eigenvalues (descending order):
        # rows nt-1, nt-2, ... of the row-major view of the column-major eigenvector matrix.
        self.vectors.copy_(self.gram[self.nt - self.rank_start:].flip(0))
        mark("evd")
        self._call("rcabeam_filtered_power_device", dev.ptr(self.signal), self.code, nv, self.nt,
                   dev.ptr(self.vectors), self.rank_start, 1, dev.ptr(self.power))
        mark("project")


def _lib_coef_16bit() -> int:
    from rcabeam_amd import _lib

    return _lib.COEF_16BIT


_C1 = {}  # the 64 + 64-element acquisition, inherited by forked CPU-baseline workers


def _c1_chunk(bounds):
    """Unfiltered OPW power of voxels [lo, hi) of the 64 + 64-element workload in the float64 oracle."""
    from oracle import combiners, das

    lo, hi = bounds
    acq, scan = _C1["acq"], _C1["scan"]
    iq = das.opw_iq(acq.rc, acq.cr, scan[lo:hi], acq.x_el, acq.y_el, acq.angles, acq.t_start, acq.c, acq.fs, acq.fd,
                    acq.fnumber)
    return combiners.power_doppler(iq)


def cpu_baseline(reps: int = 5) -> dict:
    """The float64 NumPy restatement of the reference CPU path (`_delay_numpy` -> OPW -> `power_doppler`,
    reference sim.py:344-393, opw.py:8-21, power.py:8-21; here oracle/das.py and oracle/combiners.py) on the
    64 + 64-element workload: IQ 256 x 64 x 3 x 16 per aperture, grid 32 x 64 x 32, seed 42.  Voxel blocks are
    spread over one forked worker per host core (NumPy runs each block on one thread); one warm-up run, then
    the median of `reps` runs.  Runs before the GPU is initialised, so the workers fork a GPU-free process."""
    from concurrent.futures import ProcessPoolExecutor
    from multiprocessing import get_context

    from oracle import das, scene

    acq = scene.point_scene()
    _C1["acq"], _C1["scan"] = acq, das.scan_coords(acq.grid)
    nv, nt = len(_C1["scan"]), acq.rc.shape[-1]
    cores = len(os.sched_getaffinity(0))
    workers = max(1, min(16, cores))
    edges = np.linspace(0, nv, 4 * workers + 1).astype(int)
    blocks = list(zip(edges[:-1], edges[1:]))
    times = []
    with ProcessPoolExecutor(workers, mp_context=get_context("fork")) as pool:
        for rep in range(reps + 1):
            begin = time.perf_counter()
            power = np.concatenate(list(pool.map(_c1_chunk, blocks)))
            if rep:
                times.append(time.perf_counter() - begin)
    seconds = float(np.median(times))
    _C1["power"] = power
    return {
        "value": nv * nt / seconds,
        "unit": "voxel-frames/s",
        "cores": workers,
        "kind": "port",
        "sample": f"the whole 64+64-element workload (IQ 256x64x3x16 per aperture, grid 32x64x32, seed 42), "
                  f"unfiltered OPW + power Doppler in the float64 NumPy restatement of the reference CPU path; "
                  f"{workers} forked workers on {cores} host cores, NumPy single-threaded per block; one warm-up, "
                  f"median of {reps} runs = {seconds * 1e3:.0f} ms (runs {', '.join(f'{t * 1e3:.0f}' for t in times)} ms)",
    }


def gpu_on_cpu_workload(reps: int = 5) -> dict:
    """The CPU baseline's workload on the GPU: `ensemble_pd_from_channels` with device-resident channels
    (pack + grid delay-and-sum + power, allocation included), median of `reps` after a warm-up, and its power
    against the CPU baseline's float64 power."""
    import torch

    import rcabeam_amd as rb

    acq = _C1["acq"]
    geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=acq.fnumber)
    rc, cr = (torch.from_numpy(x).cuda() for x in (acq.rc, acq.cr))
    times = []
    for rep in range(reps + 1):
        torch.cuda.synchronize()
        begin = time.perf_counter()
        pd = rb.ensemble_pd_from_channels(rc, cr, acq.angles, acq.t_start, acq.grid, geom)
        torch.cuda.synchronize()
        if rep:
            times.append(time.perf_counter() - begin)
    seconds = float(np.median(times))
    nv, nt = pd.numel(), acq.rc.shape[-1]
    got = pd.reshape(-1).cpu().numpy().astype(np.float64)
    expected = _C1["power"]
    return {"ms": seconds * 1e3, "value": nv * nt / seconds,
            "rel_l2_power_vs_cpu_baseline": float(np.linalg.norm(got - expected) / np.linalg.norm(expected))}


def host_to_host(pipe, reps: int = 3) -> dict:
    """One float32 step from host channels to a host power volume, the reference's published boundary
    (examples/_support/rca_benchmark.py:108; ROADMAP.md:45): pageable NumPy RC and CR -> pinned staged upload
    (rcabeam_upload_host) into the resident channel buffers -> the device step -> float32 power to the host.
    Timed sequentially and with the CR upload overlapping the RC pass; best of `reps` after a warm-up."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib

    hosts = [pipe.rc.cpu().numpy(), pipe.cr.cpu().numpy()]
    out = {}
    for key, overlapped in (("sequential", False), ("overlapped", True)):
        best = float("inf")
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            begin = time.perf_counter()
            if overlapped:
                pipe.step(hosts=hosts)
            else:
                for h, d in zip(hosts, (pipe.rc, pipe.cr)):
                    _lib.call("rcabeam_upload_host", h.ctypes.data, d.data_ptr(), h.nbytes, dev.upload_threads())
                pipe.step()
            power = pipe.power.to(torch.float32).cpu().numpy()
            best = min(best, time.perf_counter() - begin)
        out[key] = best
    nv, nt = pipe.g.nv, pipe.nt
    best = min(out.values())
    return {"ms": best * 1e3, "value": nv * nt / best, "unit": "voxel-frames/s",
            "sequential_ms": out["sequential"] * 1e3, "overlapped_ms": out["overlapped"] * 1e3,
            "upload_bytes": sum(h.nbytes for h in hosts), "output": f"{power.nbytes} bytes float32 power",
            "threads": dev.upload_threads(),
            "note": "sequential: both staged uploads, then the device step; overlapped: the CR upload runs while "
                    "the RC half of the delay-and-sum executes (two aperture passes)"}


def aperture_passes(pipe, reps: int = 2) -> dict:
    """The float32 delay-and-sum split into its two per-aperture grid passes (rcabeam_opw_grid_aperture_device:
    RC channels, then CR channels added to the RC sums), each timed alone with HIP events on the null stream
    (best of `reps`), beside one whole-grid launch.  Channels are the step's packed buffers."""
    import torch

    from rcabeam_amd import _device as dev

    pipe.step()  # packs RC and CR into pipe.packed
    nx, nz, ny = pipe.g.shape
    c, fs, fd, fn = pipe.g.constants
    g = pipe.g

    def one(aperture):
        pipe._call("rcabeam_opw_grid_aperture_device", dev.ptr(pipe.packed[0]), dev.ptr(pipe.packed[1]),
                   dev.ptr(g.axes[0]), dev.ptr(g.axes[1]), dev.ptr(g.axes[2]), nx, nz, ny, dev.ptr(g.xe),
                   dev.ptr(g.ye), dev.ptr(g.angles), dev.ptr(g.starts), dev.ptr(pipe.pd), dev.ptr(pipe.signal),
                   aperture, aperture, pipe.ns, pipe.nc, pipe.na, pipe.nt, c, fs, fd, fn)

    out = {}
    for key, aperture in (("rc_pass_ms", 0), ("cr_pass_ms", 1)):
        best = float("inf")
        for _ in range(reps):
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            start.record()
            one(aperture)
            end.record()
            torch.cuda.synchronize()
            best = min(best, start.elapsed_time(end))
        out[key] = best
    out["note"] = ("RC pass: the RC channels of every tile (wave-uniform aperture masks); CR pass: the CR channels "
                   "added to the RC sums, with the power")
    return out


def combiner_timing(scene, rc, cr) -> dict:
    """The combiners on the same full-size acquisition and grid: one `ensemble_pd_from_channels` call per
    method (pack + delay-and-sum + combiner epilogues; St-SW in voxel tiles bounded by its workspace), then
    clutter-filtered XDoppler and the hybrid DMAS approximation, channels resident in HBM, timed between
    synchronizations (one call each: the long kernels dominate)."""
    import torch

    from rcabeam_amd.ensemble import ensemble_pd_from_channels

    nx, nz, ny = (len(a) for a in scene.grid)
    frames = rc.shape[-1]
    out = {}
    for method, k in (("xdoppler", 2), ("rc_fmas", 2), ("dmas", 2), ("st_sw", 2), ("st_sw", 4)):
        key = method if method != "st_sw" else f"st_sw_k{k}"
        torch.cuda.synchronize()
        begin = time.perf_counter()
        ensemble_pd_from_channels(rc, cr, scene.angles, scene.t_start, scene.grid, scene.geom, method=method, k=k)
        torch.cuda.synchronize()
        seconds = time.perf_counter() - begin
        out[key] = {"ms": seconds * 1e3, "voxel_frames_per_s": nx * nz * ny * frames / seconds}
    # Clutter-filtered XDoppler (separate apertures, shared RC/CR basis, rank_start 3).
    from rcabeam_amd.ensemble import filtered_xdoppler_pd

    torch.cuda.synchronize()
    begin = time.perf_counter()
    filtered_xdoppler_pd(rc, cr, scene.angles, scene.t_start, scene.grid, scene.geom, rank_start=3)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - begin
    out["xdoppler_filtered"] = {"ms": seconds * 1e3, "voxel_frames_per_s": nx * nz * ny * frames / seconds}
    # Hybrid DMAS-CCF-ACF (delay tables + axial splines; step 0.5 samples, 16-depth slabs, Catmull-Rom
    # tables, the reference example's settings).  Its geometry assumes a 2 us acquisition start, so on this
    # scene only the timing is meaningful.  Geometry preparation is timed apart from the reconstruction.
    from rcabeam_amd.dmas_delay import prepare_hybrid

    torch.cuda.synchronize()
    begin = time.perf_counter()
    reconstruct = prepare_hybrid(tuple(rc.shape), scene.grid, scene.geom, scene.angles, 0.5, 16, cubic=True)
    torch.cuda.synchronize()
    prepare = time.perf_counter() - begin
    begin = time.perf_counter()
    reconstruct([rc, cr])
    torch.cuda.synchronize()
    seconds = time.perf_counter() - begin
    out["dmas_hybrid"] = {"ms": seconds * 1e3, "voxel_frames_per_s": nx * nz * ny * frames / seconds,
                          "prepare_ms": prepare * 1e3}
    del reconstruct
    out["dmas_hybrid_2us"] = hybrid_accuracy(scene)
    return out


def hybrid_accuracy(scene) -> dict:
    """The hybrid DMAS approximation on the same mixture re-simulated with the 2 us acquisition start its
    geometry assumes (examples/_support/dmas_delay.py:37-38, 98): its distance from the exact grid
    DMAS-CCF-ACF (relative L2 of the power volume; the reference reports 4.50 % on its own workload,
    ROADMAP.md:48) and both timings."""
    import torch

    from rcabeam_amd.dmas_delay import prepare_hybrid
    from rcabeam_amd.ensemble import ensemble_pd_from_channels
    from rcabeam_amd.scene import mixture_scene

    grid_shape = tuple(len(a) for a in scene.grid)
    s2, rc, cr = mixture_scene(grid_shape=grid_shape, t_start=2e-6)
    torch.cuda.synchronize()
    begin = time.perf_counter()
    exact = ensemble_pd_from_channels(rc, cr, s2.angles, s2.t_start, s2.grid, s2.geom, method="dmas")
    torch.cuda.synchronize()
    exact_s = time.perf_counter() - begin
    reconstruct = prepare_hybrid(tuple(rc.shape), s2.grid, s2.geom, s2.angles, 0.5, 16, cubic=True)
    reconstruct([rc, cr])
    torch.cuda.synchronize()
    begin = time.perf_counter()
    approx = reconstruct([rc, cr])
    torch.cuda.synchronize()
    hybrid_s = time.perf_counter() - begin
    a = approx.to(torch.float64).reshape(-1)
    e = exact.to(torch.float64).reshape(-1)
    return {"t_start_s": 2e-6, "exact_ms": exact_s * 1e3, "hybrid_ms": hybrid_s * 1e3,
            "rel_l2_vs_exact_dmas": float(torch.linalg.vector_norm(a - e) / torch.linalg.vector_norm(e))}


def combiner_timing_64el(reps: int = 3) -> dict:
    """Every combiner on the 64 + 64-element point-scatterer acquisition (IQ 256 x 64 x 3 x 16, seed 42)
    over a 128 x 256 x 128 grid on the same +/-6.4 mm x 5-15 mm extent: one `ensemble_pd_from_channels`
    call each, channels resident in HBM, median of `reps` after a warm-up."""
    import torch

    import rcabeam_amd as rb
    from oracle import scene as oscene

    acq = oscene.point_scene(grid_shape=(128, 256, 128))
    geom = rb.RCAGeometry(acq.x_el, acq.y_el, c=acq.c, fs=acq.fs, f_demod=acq.fd, fnumber=acq.fnumber)
    rc, cr = (torch.from_numpy(x).cuda() for x in (acq.rc, acq.cr))
    nv = 128 * 256 * 128
    frames = acq.rc.shape[-1]
    out = {}
    for method, k in (("opw", 2), ("xdoppler", 2), ("rc_fmas", 2), ("dmas", 2), ("st_sw", 2), ("st_sw", 3)):
        key = method if method != "st_sw" else f"st_sw_k{k}"
        times = []
        for rep in range(reps + 1):
            torch.cuda.synchronize()
            begin = time.perf_counter()
            rb.ensemble_pd_from_channels(rc, cr, acq.angles, acq.t_start, acq.grid, geom, method=method, k=k)
            torch.cuda.synchronize()
            if rep:
                times.append(time.perf_counter() - begin)
        seconds = float(np.median(times))
        out[key] = {"ms": seconds * 1e3, "voxel_frames_per_s": nv * frames / seconds}
    return out


def two_stage_timing(scene, rc, cr, reps: int = 2) -> dict:
    """The two-stage approximation (reference planes + not-a-knot spline expansion) on the same full-size
    acquisition and grid, channels resident in HBM: `st_sw_baseband` (one baseband plane per aperture, four
    reference depths per wavelength) and OPW IQ (eight per wavelength, all frames per pass).  Timed per
    `reconstruct` after one warm-up, workspace built beforehand; the entry point's own call (workspace
    construction included) is timed once.  The approximation's distance from the conventional forms
    (conventional St-SW k = 2, grid OPW IQ) is reported as a relative L2 error."""
    import torch

    from rcabeam_amd.ensemble import beamform_ensemble, ensemble_pd_from_channels
    from rcabeam_amd.two_stage import TwoStageOpwWorkspace, TwoStageStswWorkspace

    nx, nz, ny = (len(a) for a in scene.grid)
    frames = rc.shape[-1]
    args = (tuple(rc.shape), scene.angles, scene.t_start, scene.grid, scene.geom)

    def rel(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        begin = time.perf_counter()
        for _ in range(reps):
            result = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - begin) / reps, result

    out = {}
    stsw = TwoStageStswWorkspace(*args)
    seconds, (_, power) = timed(lambda: stsw.reconstruct(rc, cr))
    conventional = ensemble_pd_from_channels(rc, cr, scene.angles, scene.t_start, scene.grid, scene.geom,
                                             method="st_sw", k=2)
    out["st_sw_baseband"] = {"ms": seconds * 1e3, "voxel_frames_per_s": nx * nz * ny * frames / seconds,
                             "bands": len(stsw._plan.bands), "reference_depths": len(stsw.plane_depths),
                             "rel_l2_vs_conventional_st_sw": rel(power, conventional)}
    del stsw, power, conventional
    torch.cuda.synchronize()
    begin = time.perf_counter()
    ensemble_pd_from_channels(rc, cr, scene.angles, scene.t_start, scene.grid, scene.geom, method="st_sw_baseband")
    torch.cuda.synchronize()
    out["st_sw_baseband"]["entry_point_ms"] = (time.perf_counter() - begin) * 1e3
    opw = TwoStageOpwWorkspace(*args, samples_per_wavelength=8, frame_chunk=frames)
    seconds, iq = timed(lambda: opw.reconstruct(rc, cr))
    conventional = beamform_ensemble(rc, cr, scene.angles, scene.t_start, scene.grid, scene.geom)
    out["opw_iq"] = {"ms": seconds * 1e3, "voxel_frames_per_s": nx * nz * ny * frames / seconds,
                     "samples_per_wavelength": 8, "bands": len(opw._plan.bands),
                     "reference_depths": len(opw.plane_depths),
                     "rel_l2_vs_conventional_iq": rel(iq, conventional)}
    del opw, iq, conventional
    torch.cuda.empty_cache()
    return out


def rf_timing(reps: int = 3) -> dict:
    """RF -> packed IQ demodulation of one aperture at the full-size acquisition: int16 RF (3072 samples at
    60 MHz = 3 x the 20 MHz IQ rate, 128 channels, 11 angles, 200 frames, seeded), 63-tap low-pass, carrier
    15 MHz, straight into the packed (channels, angles, samples, frames) complex64 layout the delay-and-sum
    reads.  HBM roof: the RF read once plus the packed IQ written once."""
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.demod import lowpass_taps

    ns_rf, nc, na, nt, dec = 3072, 128, 11, 200, 3
    gen = torch.Generator(device="cuda").manual_seed(42)
    rf = torch.randint(-2000, 2000, (ns_rf, nc, na, nt), dtype=torch.int16, device="cuda", generator=gen)
    taps = dev.to_device(lowpass_taps(63, 8e6, 60e6), torch.float32)
    ns_iq = ns_rf // dec
    out = dev.empty((nc, na, ns_iq, nt), torch.complex64)

    def run():
        _lib.call("rcabeam_rf_to_iq_device", dev.ptr(rf), 1, ns_rf, nc, na, nt, dev.ptr(taps), 63, dec, 60e6, 15e6,
                  0.0, dev.ptr(out), ns_iq, 1)

    run()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        run()
        b.record()
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    t = float(np.median(ms))
    nbytes = rf.numel() * 2 + out.numel() * 8
    host = rf.cpu().numpy()
    torch.cuda.synchronize()
    begin = time.perf_counter()
    dev.upload(host, torch.int16)
    torch.cuda.synchronize()
    upload = time.perf_counter() - begin
    return {"workload": "one aperture: int16 RF 3072x128x11x200 at 60 MHz -> packed complex64 IQ 1024 samples, "
                        "63 taps, decimation 3", "kernel_ms": t, "bytes": nbytes, "achieved_GBps": nbytes / t / 1e6,
            "peak_GBps": HBM_PEAK_GBS, "frac": nbytes / t / 1e6 / HBM_PEAK_GBS,
            "host_rf_upload_GBps": host.nbytes / upload / 1e9, "host_rf_bytes": host.nbytes}


def ingest_timing(rc) -> dict:
    """Host-to-device upload of one aperture's complex64 channels (pageable NumPy array): torch's
    pageable copy against the library's pinned staging (rcabeam_upload_host), best of 3 each."""
    import torch

    from rcabeam_amd import _device as dev

    host = rc.cpu().numpy()
    seconds = {}
    for name, fn in (("pageable", lambda: torch.from_numpy(host).to("cuda")),
                     ("staged", lambda: dev.upload(host, torch.complex64))):
        best = float("inf")
        for _ in range(3):
            torch.cuda.synchronize()
            begin = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - begin)
            del out
        seconds[name] = best
    return {"bytes": host.nbytes, "threads": dev.upload_threads(),
            "pageable_GBps": host.nbytes / seconds["pageable"] / 1e9,
            "staged_GBps": host.nbytes / seconds["staged"] / 1e9}


def adjoint_timing(reps: int = 3) -> dict:
    """Differentiable-OPW workload: adjoint kernel time / forward kernel time.

    64 + 64 elements at 200 um, c = 1540 m/s, fs = 20 MHz, fd = 15 MHz, 3 angles (-5, 0, 5 deg),
    256 samples, 32 frames, grid 64 x 128 x 64 over +/-6.4 mm x 5-15 mm, F = 1; synthetic
    (seeded Gaussian) complex64 channels, resident in HBM; both kernels timed with HIP events.
    """
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.ensemble import _Geometry, pack_channels
    from rcabeam_amd.sim import RCAGeometry

    lib = _lib.load()
    ns, nc, na, nt = 256, 64, 3, 32
    el = ((np.arange(nc) - (nc - 1) / 2) * 200e-6).astype(np.float32)
    geom = RCAGeometry(el, el.copy(), c=1540.0, fs=20e6, f_demod=15e6, fnumber=1.0)
    grid = (np.linspace(-6.4e-3, 6.4e-3, 64), np.linspace(5e-3, 15e-3, 128), np.linspace(-6.4e-3, 6.4e-3, 64))
    angles = np.deg2rad([-5.0, 0.0, 5.0])
    g = _Geometry(angles, 0.0, grid, geom)
    gen = torch.Generator(device="cuda").manual_seed(42)
    raw = [torch.randn((ns, nc, na, nt), dtype=torch.complex64, device="cuda", generator=gen) for _ in range(2)]
    packed = [pack_channels(x, "float32") for x in raw]
    nx, nz, ny = g.shape
    c, fs, fd, fn = g.constants
    pd = dev.empty((g.nv, 2), torch.float32)
    iq = dev.empty((g.nv, nt), torch.complex64)
    grad_iq = torch.randn((g.nv, nt), dtype=torch.complex64, device="cuda", generator=gen)
    grads = [torch.zeros((nc, na, ns, nt), dtype=torch.complex64, device="cuda") for _ in range(2)]
    grad_c = torch.zeros((1,), dtype=torch.float64, device="cuda")
    axes = [dev.ptr(a) for a in g.axes]
    common = [dev.ptr(g.xe), dev.ptr(g.ye), dev.ptr(g.angles), dev.ptr(g.starts)]

    def forward():
        _lib.check(lib.rcabeam_opw_grid_device(dev.ptr(packed[0]), dev.ptr(packed[1]), 0, 0, *axes, nx, nz, ny,
                                               *common, dev.ptr(pd), dev.ptr(iq), 0, 0, ns, nc, na, nt,
                                               c, fs, fd, fn), "forward")

    def adjoint_atomic():
        _lib.check(lib.rcabeam_opw_adjoint_grid_device(dev.ptr(packed[0]), dev.ptr(packed[1]), *axes, nx, nz, ny,
                                                       *common, dev.ptr(grad_iq), dev.ptr(grads[0]),
                                                       dev.ptr(grads[1]), dev.ptr(grad_c), ns, nc, na, nt,
                                                       c, fs, fd, fn), "adjoint")

    nbytes = int(lib.rcabeam_opw_adjoint_workspace_bytes(nx, nz, ny, ns, nc, na, nt))
    work = dev.empty(((nbytes + 7) // 8,), torch.float64)

    def adjoint():
        _lib.check(lib.rcabeam_opw_adjoint_grid_deterministic_device(
            dev.ptr(packed[0]), dev.ptr(packed[1]), *axes, nx, nz, ny, *common, dev.ptr(grad_iq), dev.ptr(grads[0]),
            dev.ptr(grads[1]), dev.ptr(grad_c), dev.ptr(work), nbytes, ns, nc, na, nt, c, fs, fd, fn), "adjoint")

    times = {}
    for name, fn_ in (("forward", forward), ("adjoint", adjoint), ("adjoint_atomic", adjoint_atomic)):
        fn_()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn_()
            b.record()
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        times[name] = float(np.median(ms))
    return {
        "workload": "64+64 el, IQ 256x64x3x32, grid 64x128x64, synthetic channels",
        "forward_ms": times["forward"],
        "adjoint_ms": times["adjoint"],
        "ratio": times["adjoint"] / times["forward"],
        "adjoint": "deterministic (fixed-point gradient sums; the autograd backward)",
        "adjoint_float_atomics_ms": times["adjoint_atomic"],
    }


def time_pipeline(pipe, warmup: int, steps: int):
    """W untimed steps, then K timed steps between synchronizations: (ms per step, mean delay-and-sum ms,
    mean ms per stage).  Stage times come from HIP events recorded on the null stream between stages."""
    import torch

    for _ in range(warmup):
        pipe.step()
    torch.cuda.synchronize()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    marks = [[] for _ in range(steps)]
    log(f"bench: timing {steps} steps ({pipe.storage})")
    torch.cuda.synchronize()
    begin = time.perf_counter()
    for ev, mk in zip(events, marks):
        pipe.step(ev, mk)
    torch.cuda.synchronize()
    wall = time.perf_counter() - begin
    stages = {}
    for mk in marks:
        for (_, a), (name, b) in zip(mk[:-1], mk[1:]):
            stages.setdefault(name, []).append(a.elapsed_time(b))
    stage_ms = {name: float(np.mean(v)) for name, v in stages.items()}
    return wall * 1e3 / steps, float(np.mean([a.elapsed_time(b) for a, b in events])), stage_ms


def stage_roofs(pipe, stage_ms: dict, das_flops: float) -> dict:
    """Each stage of one step against the roof that binds it: algorithmic bytes or flops per step (inputs
    read once, outputs written once) over the stage's mean event time."""
    nv, nt, ntc = pipe.g.nv, pipe.nt, pipe.ntc
    raw = pipe.ns * pipe.nc * pipe.na * pipe.nt * 8  # complex64 channels of one aperture
    sample = 8 if pipe.code == 0 else 4  # bytes per IQ sample
    packed = pipe.ns * pipe.nc * pipe.na * ntc * (8 if pipe.storage == "float32" else 4)
    hbm = lambda nbytes, ms: {"bound": "hbm", "bytes": nbytes, "achieved_GBps": nbytes / (ms * 1e6),  # noqa: E731
                              "peak_GBps": HBM_PEAK_GBS, "frac": nbytes / (ms * 1e6) / HBM_PEAK_GBS}
    out = {}
    for name, ms in stage_ms.items():
        entry = {"ms": ms}
        if name == "channel_basis":
            rows = pipe.c_rows * 2
            entry.update(bound="mfma", note="every 4th channel row of both apertures: split-bfloat16 Gram "
                         "(4 products) + float64 EVD of the 200 x 200 channel Gram",
                         flops=4.0 * nt * (nt + 1) * rows * 4)
        elif name == "pack":
            entry.update(hbm(2 * (raw + packed), ms))
            entry["note"] = "both apertures: complex64 channels read once, packed layout written once"
        elif name == "das":
            mfma = pipe.storage != "float32"
            peak = BF16_PEAK_TFLOPS if mfma else FP32_PEAK_TFLOPS
            entry.update(bound="mfma" if mfma else "valu", flops=das_flops,
                         achieved_TFLOPs=das_flops / (ms * 1e9), peak_TFLOPs=peak,
                         frac=das_flops / (ms * 1e9) / peak,
                         note="14 flop per active tap and frame (algorithmic, not the instructions issued)")
        elif name == "expand":
            entry.update(hbm(nv * ntc * 8 + nv * nt * 8, ms))
            entry["note"] = "coded IQ read once, complex64 IQ written once"
        elif name == "gram":
            # Hermitian minimum: T (T + 1) / 2 complex multiply-adds (8 flop) per Casorati row.
            flops = 4.0 * nt * (nt + 1) * nv
            if pipe.exact_gram:
                peak, form = FP64_PEAK_TFLOPS, "exact float64 MFMA (complex64 products exact in float64)"
            elif pipe.code == 0:
                peak, form = BF16_PEAK_TFLOPS, "split bfloat16 MFMA, 4 products per complex64 product"
                flops *= 4
            else:
                peak, form = BF16_PEAK_TFLOPS, "one exact 16-bit MFMA product"
            entry.update(bound="mfma", flops=flops, achieved_TFLOPs=flops / (ms * 1e9), peak_TFLOPs=peak,
                         frac=flops / (ms * 1e9) / peak, bytes=nv * nt * sample,
                         note=f"4 (T + 1) flop per voxel-frame (upper triangle of the complex Gram); {form}"
                         + ("" if pipe.exact_gram else "; includes the |max| pass"))
        elif name == "evd":
            entry.update(bound="latency", note="200 x 200 complex128 Hermitian EVD (one-workgroup Householder "
                         "tridiagonalization + rocSOLVER zstedc) and the T x T Hermitian assembly")
        elif name == "project":
            entry.update(hbm(nv * nt * sample + nv * 8, ms))
            entry["note"] = "IQ read once, float64 filtered power written once (3 discarded vectors)"
        out[name] = entry
    return out


def traffic_from_profiles(precision: str):
    """HBM bytes per delay-and-sum launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/das_grid_hbm_traffic.json), already scaled by the copy-kernel calibration; None if absent."""
    summary = PROFILES / "das_grid_hbm_traffic.json"
    if not summary.exists():
        return None
    try:
        entry = json.loads(summary.read_text()).get(precision)
        return None if entry is None else float(entry["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError, TypeError):
        return None


def sampled_oracle(scene, rc, cr, pick):
    """Float64 oracle OPW IQ at the voxels `pick` of the full-size workload (parity lines of the record)."""
    from oracle import das

    scan = das.scan_coords(scene.grid)
    g = scene.geom
    return das.opw_iq(rc.cpu().numpy(), cr.cpu().numpy(), scan[pick], g.x_el, g.y_el, scene.angles, scene.t_start,
                      g.c, g.fs, g.f_demod, g.fnumber)


def main() -> int:
    args = parse()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); multi-GPU runs are not part of "
              "this build (the path would shard by voxel slabs, with one all-reduce of the T x T Gram)",
              file=sys.stderr)
        return 2
    baseline = None
    if not (args.no_cpu_baseline or args.two_stage_only or args.combiners_only or args.aperture_passes_only):
        log("bench: CPU baseline (float64 NumPy restatement, 64+64-element workload)")
        baseline = cpu_baseline()
    import torch

    from rcabeam_amd import _device as dev
    from rcabeam_amd import _lib
    from rcabeam_amd.scene import mixture_scene

    lib = _lib.load()
    rank_start = 3
    log(f"bench: generating scene grid={tuple(args.grid)}")
    scene, rc, cr = mixture_scene(grid_shape=tuple(args.grid))
    if args.two_stage_only:
        print(json.dumps({"two_stage": two_stage_timing(scene, rc, cr)}), flush=True)
        return 0
    if args.combiners_only:
        print(json.dumps({"combiners": combiner_timing(scene, rc, cr)}), flush=True)
        return 0
    storage = args.precision
    pipe = Pipeline(scene, rc, cr, storage, rank_start, args.lowrank_modes)
    if args.aperture_passes_only:
        if storage != "float32":
            log("bench: --aperture-passes-only times the float32 form")
            return 2
        print(json.dumps({"das_aperture_passes": aperture_passes(pipe)}), flush=True)
        return 0
    nv, nt = pipe.g.nv, pipe.nt
    counter = dev.empty((1,), torch.int64)
    counter.zero_()
    c, fs, fd, fn = pipe.g.constants
    _lib.check(lib.rcabeam_count_active_taps_device(dev.ptr(pipe.g.scan), dev.ptr(pipe.g.xe), dev.ptr(pipe.g.ye),
                                                    dev.ptr(pipe.g.angles), dev.ptr(pipe.g.starts), pipe.ns, pipe.nc,
                                                    pipe.na, nv, c, fs, fd, fn, dev.ptr(counter)), "count taps")
    taps = int(counter.item())
    das_flops = FLOPS_PER_TAP * taps * nt
    log(f"bench: {taps} active taps ({taps / nv:.1f} per voxel), warmup {args.warmup}")
    ms_per_step, das_ms, stage_ms = time_pipeline(pipe, args.warmup, args.steps)
    value = nv * nt / (ms_per_step / 1e3)
    achieved = das_flops / (das_ms / 1e3) / 1e12
    mfma_das = storage != "float32"
    peak = BF16_PEAK_TFLOPS if mfma_das else FP32_PEAK_TFLOPS
    # Clutter cut: singular values sigma_k = sqrt(lambda_k) of the Gram, descending; the retained interval
    # starts at rank_start, so sigma_{rank_start} / sigma_{rank_start + 1} (one-based) is the margin of the cut.
    lam = pipe.values.cpu().numpy()[::-1]
    sigma = np.sqrt(np.maximum(lam, 0))
    record = {
        "metric": "voxel-frames/s OPW+SVD+PD",
        "value": value,
        "unit": "voxel-frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"float32": "f32", "float16": "f16-storage/f32-accumulate", "bfloat16": "bf16-storage/f32-accumulate",
                  "float16-lowrank": "f16-residual-storage/f32-accumulate",
                  "bfloat16-lowrank": "bf16-residual-storage/f32-accumulate"}[storage],
        "data": "synthetic (GPU-simulated tissue + blood + noise scene, seed 42)",
        "config": {
            "workload": "filtered OPW PD: 128+128 el, IQ 1024x128x11x200 per aperture, grid "
                        f"{'x'.join(map(str, args.grid))}, rank_start {rank_start}",
            "precision": args.precision,
            "das_ms": das_ms,
            "active_taps_per_voxel": taps / nv,
        },
        "roofline": {
            "bound": "mfma" if mfma_das else "valu",
            "engine": ("dense fp16/bf16 MFMA" if mfma_das else
                       "fp32 VALU (MI355X fp32 vector peak; the build disables packed FP32, so half of it is reachable)"),
            "kernel": DAS_KERNEL[storage],
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic_from_profiles(storage),
        },
        "stages": stage_roofs(pipe, stage_ms, das_flops),
        "clutter_cut": {"rank_start": rank_start, "sigma_1_to_6": sigma[:6].tolist(),
                        "gap_sigma3_over_sigma4": float(sigma[rank_start - 1] / sigma[rank_start])},
    }
    # Voxels at which every form's IQ is compared with the float64 oracle.
    pick = np.sort(np.random.default_rng(0).choice(nv, size=args.cpu_voxels, replace=False))
    pick_d = torch.from_numpy(pick).cuda()
    sampled = {"float32": pipe.signal[pick_d].clone() if pipe.code == 0 else None}
    if storage == "float32" and not args.no_mixed:
        # Mixed-precision forms on the same scene, with their full-size difference from the float32 pipeline.
        reference = {"iq": pipe.signal, "pd": pipe.pd[:, 0].clone(), "power": pipe.power.clone()}

        def rel(a, b):
            a = a.to(torch.complex128) if a.is_complex() else a.to(torch.float64)
            b = b.to(torch.complex128) if b.is_complex() else b.to(torch.float64)
            return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))

        for key, form, label in (
            ("mixed", "float16-lowrank",
             f"float16 channels coded as {args.lowrank_modes} temporal modes (16-bit hi + lo) + 16-bit residual, "
             "MFMA delay-and-sum with one 16-bit coefficient per sample, float32 accumulation, complex64 IQ"),
            ("mixed_bf16", "bfloat16-lowrank",
             f"bfloat16 channels coded as {args.lowrank_modes} temporal modes (16-bit hi + lo) + 16-bit residual, "
             "MFMA delay-and-sum with one 16-bit coefficient per sample, float32 accumulation, complex64 IQ"),
            ("mixed_direct", "float16",
             "float16 channels and IQ rounded directly, MFMA delay-and-sum with one 16-bit coefficient per sample"),
        ):
            log(f"bench: mixed form {form}")
            mixed = Pipeline(scene, rc, cr, form, rank_start, args.lowrank_modes)
            m_ms, m_das, m_stages = time_pipeline(mixed, args.warmup, args.steps)
            iq = mixed.signal if mixed.code == 0 else mixed.signal.view(torch.float16).to(torch.float32).view(
                torch.complex64)
            record[key] = {
                "precision": label,
                "value": nv * nt / (m_ms / 1e3),
                "ms_per_step": m_ms,
                "das_ms": m_das,
                "das_tflops_algorithmic": das_flops / (m_das / 1e3) / 1e12,
                "speedup_vs_float32": ms_per_step / m_ms,
                "rel_l2_vs_float32": {"iq": rel(iq, reference["iq"]), "power": rel(mixed.pd[:, 0], reference["pd"]),
                                      "filtered_power": rel(mixed.power, reference["power"])},
                "stages_ms": m_stages,
            }
            sampled[key] = iq[pick_d].clone()
            del mixed, iq
            torch.cuda.empty_cache()
        del reference
    if storage == "float32":
        log("bench: delay-and-sum as two per-aperture passes")
        record["das_aperture_passes"] = aperture_passes(pipe)
    log("bench: host to host (staged upload + step + power download)")
    record["host_to_host"] = host_to_host(pipe)
    log("bench: adjoint / forward on the differentiable-OPW workload")
    record["adjoint"] = adjoint_timing()
    if not args.no_combiners:
        log("bench: combiners at full size")
        record["combiners"] = combiner_timing(scene, rc, cr)
        log("bench: combiners on the 64+64-element acquisition at 128x256x128")
        record["combiners_64el"] = combiner_timing_64el()
        log("bench: host ingest")
        record["host_ingest"] = ingest_timing(rc)
    log("bench: RF to IQ demodulation")
    record["rf_to_iq"] = rf_timing()
    if not args.no_two_stage:
        log("bench: two-stage approximation at full size")
        record["two_stage"] = two_stage_timing(scene, rc, cr)
    if baseline is not None:
        baseline["gpu_same_workload"] = gpu_on_cpu_workload()
        record["cpu_baseline"] = baseline
    if args.cpu_voxels:
        log(f"bench: float64 oracle at {args.cpu_voxels} sampled voxels")
        oracle_iq = sampled_oracle(scene, rc, cr, pick)
        # Each form's IQ and unfiltered power at the sampled voxels against the float64 oracle.
        for key, iq in sampled.items():
            if iq is None:
                continue
            got = iq.cpu().numpy().astype(np.complex128)
            target = record if key == "float32" else record[key]
            target["vs_float64_oracle_sampled"] = {
                "voxels": len(pick),
                "iq": float(np.linalg.norm(got - oracle_iq) / np.linalg.norm(oracle_iq)),
                "power": float(np.linalg.norm((np.abs(got) ** 2).mean(-1) - (np.abs(oracle_iq) ** 2).mean(-1))
                               / np.linalg.norm((np.abs(oracle_iq) ** 2).mean(-1))),
            }
    print(json.dumps(record), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
