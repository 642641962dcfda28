"""numpy-facing wrappers over the C ABI (used by tests, bench and the server).

Arrays cross the boundary as FP64 numpy arrays in the reference layouts:
2D index i*nz+j, 3D index (i*ny2+j)*nz2+k, z fastest.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import check, lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _sp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_size_t))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def ricker_length(f: float, dt: float) -> int:
    return int(lib.sm_ricker_length(f, dt))


def source_functions(A, f, dt: float) -> tuple[np.ndarray, np.ndarray]:
    """compute_src_fct (src.c:43-65): returns (lengths, concatenated fct)."""
    A = _f64(np.atleast_1d(A))
    f = _f64(np.atleast_1d(f))
    lengths = np.array([ricker_length(fi, dt) for fi in f], dtype=np.uintp)
    fct = np.zeros(int(lengths.sum()), dtype=np.float64)
    check(lib.sm_compute_src_fct(len(A), _dp(A), _dp(f), dt, _sp(lengths), _dp(fct)),
          "sm_compute_src_fct")
    return lengths, fct


def nsteps(duration: float, dt: float) -> int:
    return int(lib.sm_nsteps(duration, dt))


def record_due(it: int, dt: float, t0: float, rdt: float) -> bool:
    return bool(lib.sm_record_due(it, dt, t0, rdt))


_CONTEXTS = weakref.WeakSet()


@atexit.register
def _close_all() -> None:
    """Destroy live contexts (and their solvers) while the HIP runtime is
    still fully up, rather than from interpreter teardown."""
    for c in list(_CONTEXTS):
        try:
            c.close()
        except Exception:
            pass


class Context:
    """One device + one HIP stream."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib.sm_ctx_create(device, C.byref(h)), "sm_ctx_create")
        self.handle = h
        # solvers and batches on this context; closed before the context so
        # that no destroy call ever sees a freed stream
        self._children = weakref.WeakSet()
        _CONTEXTS.add(self)

    def adopt(self, child) -> None:
        self._children.add(child)

    def sync(self) -> None:
        check(lib.sm_ctx_sync(self.handle), "sm_ctx_sync")

    def mem_info(self) -> tuple[int, int]:
        free, total = C.c_size_t(), C.c_size_t()
        check(lib.sm_ctx_mem_info(self.handle, C.byref(free), C.byref(total)), "sm_ctx_mem_info")
        return free.value, total.value

    def probe_stream(self, n: int, reads: int, writes: int, iters: int = 20, vec: int = 1) -> float:
        """ms per launch of a plain kernel streaming reads+writes FP32 arrays."""
        ms = C.c_double()
        check(lib.sm_probe_stream(self.handle, n, reads, writes, vec, iters, C.byref(ms)),
              "sm_probe_stream")
        return ms.value

    def close(self) -> None:
        if self.handle:
            for c in list(self._children):
                c.close()
            check(lib.sm_ctx_destroy(self.handle), "sm_ctx_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_STORAGE = {"f32": L.SM_STORE_F32, "f64": L.SM_STORE_F64, "bf16": L.SM_STORE_BF16,
            "f16": L.SM_STORE_F16}


@dataclass
class A3dConfig:
    nx: int
    ny: int
    nz: int
    np: int
    dx: float
    dy: float
    dz: float
    dt: float
    f0: float
    vmax: float = 5000.0
    pml_order: float = 2.0
    Rc: float = 1e-5
    kappa_max: float = 1.0
    alpha_max: int = 1
    storage: str = "f32"
    naive: bool = False
    sround: bool = False      # BF16/FP16: stochastic rounding of stored values
    pad64: bool = False       # rows always padded to 64 cells; else when <= 5% extra
    mat16: bool = False       # BF16/FP16: material as FP16 with power-of-two scales (tile variants)

    def params(self) -> L.A3dParams:
        return L.A3dParams(self.nx, self.ny, self.nz, self.np, self.dx, self.dy, self.dz,
                           self.dt, self.vmax, self.pml_order, self.Rc, self.kappa_max,
                           self.alpha_max, self.f0, _STORAGE[self.storage],
                           (L.SM_A3D_NAIVE if self.naive else 0)
                           | (L.SM_A3D_SROUND if self.sround else 0)
                           | (L.SM_A3D_PAD64 if self.pad64 else 0)
                           | (L.SM_A3D_MAT16 if self.mat16 else 0))

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.nx + 2 * self.np, self.ny + 2 * self.np, self.nz + 2 * self.np)

    def bytes_needed(self, nrec: int = 0, max_steps: int = 0) -> int:
        """Device bytes a solver with this configuration will hold."""
        p = self.params()
        return int(lib.sm_a3d_bytes_needed(C.byref(p), nrec, max_steps))


class Acoustic3D:
    """3D acoustic CPML solver (a_iso_3d)."""

    def __init__(self, ctx: Context, cfg: A3dConfig, vp: np.ndarray, rho: np.ndarray):
        self.ctx, self.cfg = ctx, cfg
        n = int(np.prod(cfg.shape))
        self._vp, self._rho = _f64(vp).reshape(-1), _f64(rho).reshape(-1)
        if self._vp.size != n or self._rho.size != n:
            raise ValueError(f"model must hold {n} padded cells")
        self._p = cfg.params()
        h = C.c_void_p()
        ctx.adopt(self)
        check(lib.sm_a3d_create(ctx.handle, C.byref(self._p), _dp(self._vp), _dp(self._rho),
                                C.byref(h)), "sm_a3d_create")
        self.handle = h
        self.nrec = 0

    def set_sources(self, idx, lengths, fct) -> None:
        idx = np.ascontiguousarray(idx, dtype=np.uintp)
        lengths = np.ascontiguousarray(lengths, dtype=np.uintp)
        fct = _f64(fct)
        check(lib.sm_a3d_set_sources(self.handle, len(idx), _sp(idx), _sp(lengths), _dp(fct)),
              "sm_a3d_set_sources")

    def set_receivers(self, idx, comp, max_steps: int) -> None:
        idx = np.ascontiguousarray(idx, dtype=np.uintp)
        comp = np.ascontiguousarray(comp, dtype=np.intc)
        check(lib.sm_a3d_set_receivers(self.handle, len(idx), _sp(idx), _ip(comp), max_steps),
              "sm_a3d_set_receivers")
        self.nrec = len(idx)

    def run(self, nsteps: int) -> None:
        check(lib.sm_a3d_run(self.handle, nsteps), "sm_a3d_run")

    @property
    def step(self) -> int:
        it = C.c_size_t()
        check(lib.sm_a3d_current_step(self.handle, C.byref(it)), "sm_a3d_current_step")
        return it.value

    def traces(self, it0: int, count: int) -> np.ndarray:
        out = np.zeros((count, self.nrec), dtype=np.float64)
        check(lib.sm_a3d_get_traces(self.handle, it0, count, _dp(out)), "sm_a3d_get_traces")
        return out

    def field(self, comp: int) -> np.ndarray:
        out = np.zeros(self.cfg.shape, dtype=np.float64)
        check(lib.sm_a3d_get_field(self.handle, comp, _dp(out)), "sm_a3d_get_field")
        return out

    def slice_xz(self, comp: int, iy: int) -> np.ndarray:
        nx2, _, nz2 = self.cfg.shape
        out = np.zeros((nx2, nz2), dtype=np.float64)
        check(lib.sm_a3d_get_slice_xz(self.handle, comp, iy, _dp(out)), "sm_a3d_get_slice_xz")
        return out

    def reset(self) -> None:
        check(lib.sm_a3d_reset(self.handle), "sm_a3d_reset")

    def set_launch(self, rows_per_block: int, x_chunks: int) -> None:
        check(lib.sm_a3d_set_launch(self.handle, rows_per_block, x_chunks), "sm_a3d_set_launch")

    def launch(self) -> tuple[int, int]:
        tj, ch = C.c_int(), C.c_int()
        check(lib.sm_a3d_get_launch(self.handle, C.byref(tj), C.byref(ch)), "sm_a3d_get_launch")
        return tj.value, ch.value

    VARIANTS = {"cell": L.SM_A3D_CELL, "cell2": L.SM_A3D_CELL2, "tile": L.SM_A3D_TILE,
                "tilesplit": L.SM_A3D_TILE_SPLIT, "tile_nt": L.SM_A3D_TILE_NT,
                "tilesplit_nt": L.SM_A3D_TILE_SPLIT_NT}

    def set_variant(self, variant: str) -> None:
        """Fused kernel variant: "cell" (one cell per lane, 63-cell z tiles),
        "cell2" (the same, loading two planes ahead), "tile" (64-cell z tiles
        aligned to the device rows, staged through LDS, 16-byte stores),
        "tilesplit" (the same with the CPML work compiled in per tile only
        where a layer is touched), "tile_nt"/"tilesplit_nt" (those two with
        non-temporal state stores; rows_per_block 4 or 8); results are
        bit-identical."""
        check(lib.sm_a3d_set_variant(self.handle, self.VARIANTS[variant]), "sm_a3d_set_variant")

    def variant(self) -> str:
        v = C.c_int()
        check(lib.sm_a3d_get_variant(self.handle, C.byref(v)), "sm_a3d_get_variant")
        return {n: k for k, n in self.VARIANTS.items()}.get(v.value, "none")

    def set_timing(self, on: bool) -> None:
        check(lib.sm_a3d_set_timing(self.handle, 1 if on else 0), "sm_a3d_set_timing")

    def timing(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_size_t()
        check(lib.sm_a3d_get_timing(self.handle, C.byref(ms), C.byref(n)), "sm_a3d_get_timing")
        return ms.value, n.value

    def device_bytes(self) -> int:
        b = C.c_size_t()
        check(lib.sm_a3d_device_bytes(self.handle, C.byref(b)), "sm_a3d_device_bytes")
        return b.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib.sm_a3d_destroy(self.handle), "sm_a3d_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class E2dConfig:
    nx: int
    nz: int
    npml: int
    h: float
    dt: float
    storage: str = "f32"
    q1_compat: bool = True
    damping_scale: float = 1.0
    naive: bool = False       # FP32 with the two reference-order kernels

    def params(self) -> L.E2dParams:
        return L.E2dParams(self.nx, self.nz, self.npml, self.h, self.dt, _STORAGE[self.storage],
                           1 if self.q1_compat else 0, self.damping_scale,
                           1 if self.naive else 0)

    @property
    def padded(self) -> tuple[int, int]:
        return (self.nx + 2 * self.npml, self.nz + 2 * self.npml)

    def bytes_needed(self, nrec: int = 0, max_steps: int = 0) -> int:
        """Device bytes a solver with this configuration will hold."""
        p = self.params()
        return int(lib.sm_e2d_bytes_needed(C.byref(p), nrec, max_steps))


class Elastic2D:
    """2D elastic split-PML solver (e_iso)."""

    def __init__(self, ctx: Context, cfg: E2dConfig, vp, vs, rho):
        self.ctx, self.cfg = ctx, cfg
        n = cfg.nx * cfg.nz
        self._m = [_f64(a).reshape(-1) for a in (vp, vs, rho)]
        if any(a.size != n for a in self._m):
            raise ValueError(f"model must hold {n} cells")
        self._p = cfg.params()
        h = C.c_void_p()
        ctx.adopt(self)
        check(lib.sm_e2d_create(ctx.handle, C.byref(self._p), *(_dp(a) for a in self._m),
                                C.byref(h)), "sm_e2d_create")
        self.handle = h
        self.nrec = 0

    def set_sources(self, types, idx, lengths, fct) -> None:
        types = np.ascontiguousarray(types, dtype=np.intc)
        idx = np.ascontiguousarray(idx, dtype=np.uintp)
        lengths = np.ascontiguousarray(lengths, dtype=np.uintp)
        fct = _f64(fct)
        check(lib.sm_e2d_set_sources(self.handle, len(idx), _ip(types), _sp(idx), _sp(lengths),
                                     _dp(fct)), "sm_e2d_set_sources")

    def set_receivers(self, idx, comp, max_steps: int) -> None:
        idx = np.ascontiguousarray(idx, dtype=np.uintp)
        comp = np.ascontiguousarray(comp, dtype=np.intc)
        check(lib.sm_e2d_set_receivers(self.handle, len(idx), _sp(idx), _ip(comp), max_steps),
              "sm_e2d_set_receivers")
        self.nrec = len(idx)

    def run(self, nsteps: int) -> None:
        check(lib.sm_e2d_run(self.handle, nsteps), "sm_e2d_run")

    @property
    def step(self) -> int:
        it = C.c_size_t()
        check(lib.sm_e2d_current_step(self.handle, C.byref(it)), "sm_e2d_current_step")
        return it.value

    def traces(self, it0: int, count: int) -> np.ndarray:
        out = np.zeros((count, self.nrec), dtype=np.float64)
        check(lib.sm_e2d_get_traces(self.handle, it0, count, _dp(out)), "sm_e2d_get_traces")
        return out

    def field(self, comp: int) -> np.ndarray:
        out = np.zeros((self.cfg.nx, self.cfg.nz), dtype=np.float64)
        check(lib.sm_e2d_get_field(self.handle, comp, _dp(out)), "sm_e2d_get_field")
        return out

    def field_padded(self, comp: int) -> np.ndarray:
        out = np.zeros(self.cfg.padded, dtype=np.float64)
        check(lib.sm_e2d_get_field_padded(self.handle, comp, _dp(out)), "sm_e2d_get_field_padded")
        return out

    def reset(self) -> None:
        check(lib.sm_e2d_reset(self.handle), "sm_e2d_reset")

    def set_launch(self, rows_per_chunk: int, graph_steps: int = 32) -> None:
        check(lib.sm_e2d_set_launch(self.handle, rows_per_chunk, graph_steps), "sm_e2d_set_launch")

    def launch(self) -> tuple[int, int]:
        a, b = C.c_int(), C.c_int()
        check(lib.sm_e2d_get_launch(self.handle, C.byref(a), C.byref(b)), "sm_e2d_get_launch")
        return a.value, b.value

    E2D_VARIANTS = {"rows": L.SM_E2D_ROWS, "staged": L.SM_E2D_STAGED}

    def set_variant(self, variant: str) -> None:
        """Kernel form of the fused FP32 step: "rows" or "staged" (identical results)."""
        check(lib.sm_e2d_set_variant(self.handle, self.E2D_VARIANTS[variant]), "sm_e2d_set_variant")

    def variant(self) -> str:
        v = C.c_int()
        check(lib.sm_e2d_get_variant(self.handle, C.byref(v)), "sm_e2d_get_variant")
        return {n: k for k, n in self.E2D_VARIANTS.items()}.get(v.value, "none")

    def set_timing(self, on: bool) -> None:
        check(lib.sm_e2d_set_timing(self.handle, 1 if on else 0), "sm_e2d_set_timing")

    def timing(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_size_t()
        check(lib.sm_e2d_get_timing(self.handle, C.byref(ms), C.byref(n)), "sm_e2d_get_timing")
        return ms.value, n.value

    def device_bytes(self) -> int:
        b = C.c_size_t()
        check(lib.sm_e2d_device_bytes(self.handle, C.byref(b)), "sm_e2d_device_bytes")
        return b.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib.sm_e2d_destroy(self.handle), "sm_e2d_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class E2dBatch:
    """Runs many Elastic2D shots (FP32 fused) together: one launch per time
    step over all shots still running, replayed from hipGraphs."""

    def __init__(self, ctx: Context, graph_steps: int = 32):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.adopt(self)
        check(lib.sm_e2d_batch_create(ctx.handle, C.byref(h)), "sm_e2d_batch_create")
        self.handle = h
        check(lib.sm_e2d_batch_set_graph_steps(h, graph_steps), "sm_e2d_batch_set_graph_steps")

    def run(self, shots, nsteps) -> None:
        arr = (C.c_void_p * len(shots))(*[s.handle.value for s in shots])
        steps = np.ascontiguousarray(nsteps, dtype=np.uintp)
        if steps.size != len(shots):
            raise ValueError("one step count per shot")
        check(lib.sm_e2d_batch_run(self.handle, len(shots), arr, _sp(steps)), "sm_e2d_batch_run")

    def set_group_bytes(self, nbytes: int) -> None:
        """Run the shots in groups whose wavefield/material memory stays
        within nbytes, one group after another (0: all shots in every step
        launch).  Identical results; a group that fits the last-level cache
        keeps its fields there between steps."""
        check(lib.sm_e2d_batch_set_group_bytes(self.handle, int(nbytes)),
              "sm_e2d_batch_set_group_bytes")

    def set_variant(self, variant: str) -> None:
        """Kernel form for the batch: "rows" or "staged" (identical results)."""
        check(lib.sm_e2d_batch_set_variant(self.handle, Elastic2D.E2D_VARIANTS[variant]),
              "sm_e2d_batch_set_variant")

    def set_timing(self, on: bool) -> None:
        check(lib.sm_e2d_batch_set_timing(self.handle, 1 if on else 0), "sm_e2d_batch_set_timing")

    def timing(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_size_t()
        check(lib.sm_e2d_batch_get_timing(self.handle, C.byref(ms), C.byref(n)),
              "sm_e2d_batch_get_timing")
        return ms.value, n.value

    def graphs(self) -> tuple[int, int]:
        """(captured graphs held now, graphs instantiated since creation)."""
        live, built = C.c_size_t(), C.c_size_t()
        check(lib.sm_e2d_batch_get_graphs(self.handle, C.byref(live), C.byref(built)),
              "sm_e2d_batch_get_graphs")
        return live.value, built.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib.sm_e2d_batch_destroy(self.handle), "sm_e2d_batch_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class CylConfig:
    """e_cyl_iso grid: nx radial and nz vertical interior nodes, np CPML
    layers on the outer r side and both z ends (padded nx+np by nz+2np)."""
    nx: int
    nz: int
    np: int
    dr: float
    dz: float
    dt: float
    f0: float
    mode: int = 0             # azimuthal mode n
    vmax: float = 5000.0
    pml_order: float = 2.0
    Rc: float = 1e-5
    kappa_max: float = 1.0
    alpha_max: int = 1
    storage: str = "f64"

    def params(self) -> L.CylParams:
        return L.CylParams(self.nx, self.nz, self.np, self.dr, self.dz, self.dt, self.mode,
                           self.vmax, self.pml_order, self.Rc, self.kappa_max, self.alpha_max,
                           self.f0, _STORAGE[self.storage], 0)

    @property
    def shape(self) -> tuple[int, int]:
        return (self.nx + self.np, self.nz + 2 * self.np)

    def bytes_needed(self, nrec: int = 0, max_steps: int = 0) -> int:
        p = self.params()
        return int(lib.sm_cyl_bytes_needed(C.byref(p), nrec, max_steps))


def kurkjian_fct(lam: float, dr: float, dz: float, r: float, mode: int, A: float, f: float,
                 dt: float) -> np.ndarray:
    """compute_kurkjian (src.c:202-229) for a source with Lame constant lam
    at radius r."""
    n = ricker_length(f, dt)
    out = np.zeros(n)
    check(lib.sm_kurkjian_fct(lam, dr, dz, r, mode, A, f, dt, n, _dp(out)), "sm_kurkjian_fct")
    return out


class Cylindrical:
    """2.5D elastic solver in cylindrical coordinates (e_cyl_iso)."""

    FIELDS = {"vr": L.SM_VR, "vt": L.SM_VT, "vz": L.SM_VZ, "trr": L.SM_TRR, "ttt": L.SM_TTT,
              "tzz": L.SM_TZZ, "trz": L.SM_TRZ, "trt": L.SM_TRT, "ttz": L.SM_TTZ}

    def __init__(self, ctx: Context, cfg: CylConfig, vp, vs, rho):
        self.ctx, self.cfg = ctx, cfg
        n = int(np.prod(cfg.shape))
        self._m = [_f64(a).reshape(-1) for a in (vp, vs, rho)]
        if any(a.size != n for a in self._m):
            raise ValueError(f"model must hold {n} padded cells")
        self._p = cfg.params()
        h = C.c_void_p()
        ctx.adopt(self)
        check(lib.sm_cyl_create(ctx.handle, C.byref(self._p), *(_dp(a) for a in self._m),
                                C.byref(h)), "sm_cyl_create")
        self.handle = h
        self.nrec = 0

    def set_sources(self, types, idx, lengths, fct) -> None:
        types = np.ascontiguousarray(types, dtype=np.intc)
        idx = np.ascontiguousarray(idx, dtype=np.uintp)
        lengths = np.ascontiguousarray(lengths, dtype=np.uintp)
        fct = _f64(fct)
        check(lib.sm_cyl_set_sources(self.handle, len(idx), _ip(types), _sp(idx), _sp(lengths),
                                     _dp(fct)), "sm_cyl_set_sources")

    def set_receivers(self, idx, comp, max_steps: int) -> None:
        idx = np.ascontiguousarray(idx, dtype=np.uintp)
        comp = np.ascontiguousarray(comp, dtype=np.intc)
        check(lib.sm_cyl_set_receivers(self.handle, len(idx), _sp(idx), _ip(comp), max_steps),
              "sm_cyl_set_receivers")
        self.nrec = len(idx)

    def run(self, nsteps: int) -> None:
        check(lib.sm_cyl_run(self.handle, nsteps), "sm_cyl_run")

    def set_timing(self, on: bool) -> None:
        """HIP events around the two update launches of every step."""
        check(lib.sm_cyl_set_timing(self.handle, 1 if on else 0), "sm_cyl_set_timing")

    def timing(self) -> tuple[float, int]:
        """(milliseconds, steps) timed since set_timing(True)."""
        ms, n = C.c_double(), C.c_size_t()
        check(lib.sm_cyl_get_timing(self.handle, C.byref(ms), C.byref(n)), "sm_cyl_get_timing")
        return ms.value, n.value

    @property
    def step(self) -> int:
        it = C.c_size_t()
        check(lib.sm_cyl_current_step(self.handle, C.byref(it)), "sm_cyl_current_step")
        return it.value

    def traces(self, it0: int, count: int) -> np.ndarray:
        out = np.zeros((count, self.nrec), dtype=np.float64)
        check(lib.sm_cyl_get_traces(self.handle, it0, count, _dp(out)), "sm_cyl_get_traces")
        return out

    def field(self, name: str) -> np.ndarray:
        out = np.zeros(self.cfg.shape, dtype=np.float64)
        check(lib.sm_cyl_get_field(self.handle, self.FIELDS[name], _dp(out)), "sm_cyl_get_field")
        return out

    def reset(self) -> None:
        check(lib.sm_cyl_reset(self.handle), "sm_cyl_reset")

    def device_bytes(self) -> int:
        b = C.c_size_t()
        check(lib.sm_cyl_device_bytes(self.handle, C.byref(b)), "sm_cyl_device_bytes")
        return b.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib.sm_cyl_destroy(self.handle), "sm_cyl_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ShConfig:
    """ve_vti_sh_pml grid: nx by nz interior nodes and np CPML layers on every
    side (none when abs is False), padded (nx + 2np) by (nz + 2np)."""
    nx: int
    nz: int
    np: int
    dx: float
    dz: float
    dt: float
    f0: float
    vmax: float = 5000.0
    pml_order: float = 2.0
    Rc: float = 1e-5
    kappa_max: float = 1.0
    alpha_max: int = 1
    abs: bool = True
    storage: str = "f64"
    hipfft: bool = False      # every derivative through hipFFT (SM_PSTD_HIPFFT)

    def params(self) -> L.ShParams:
        return L.ShParams(self.nx, self.nz, self.np, self.dx, self.dz, self.dt, self.vmax,
                          self.pml_order, self.Rc, self.kappa_max, self.alpha_max, self.f0,
                          1 if self.abs else 0, _STORAGE[self.storage], L.SM_PSTD_HIPFFT if self.hipfft else 0)

    @property
    def shape(self) -> tuple[int, int]:
        np_ = self.np if self.abs else 0
        return (self.nx + 2 * np_, self.nz + 2 * np_)

    def bytes_needed(self, nrec: int = 0, max_steps: int = 0) -> int:
        p = self.params()
        return int(lib.sm_sh_bytes_needed(C.byref(p), nrec, max_steps))


class _Pstd:
    """Common driver of the pseudo-spectral solvers: `_pre` is the C prefix
    (sm_sh, sm_ve), FIELDS the recordable components."""

    _pre = ""
    FIELDS: dict = {}

    def _f(self, name):
        return getattr(lib, f"{self._pre}_{name}")

    def set_sources(self, types, idx, lengths, fct) -> None:
        t = np.ascontiguousarray(types, dtype=np.intc)
        i = np.ascontiguousarray(idx, dtype=np.uintp)
        ln = np.ascontiguousarray(lengths, dtype=np.uintp)
        f = _f64(fct)
        check(self._f("set_sources")(self.handle, len(t), t.ctypes.data_as(C.POINTER(C.c_int)), _sp(i),
                                     _sp(ln), _dp(f)), f"{self._pre}_set_sources")

    def set_receivers(self, idx, comp, max_steps: int) -> None:
        i = np.ascontiguousarray(idx, dtype=np.uintp)
        c = np.ascontiguousarray(comp, dtype=np.intc)
        check(self._f("set_receivers")(self.handle, len(i), _sp(i), c.ctypes.data_as(C.POINTER(C.c_int)),
                                       max_steps), f"{self._pre}_set_receivers")
        self.nrec = len(i)

    def run(self, nsteps: int) -> None:
        check(self._f("run")(self.handle, nsteps), f"{self._pre}_run")

    def set_timing(self, on: bool) -> None:
        check(self._f("set_timing")(self.handle, 1 if on else 0), f"{self._pre}_set_timing")

    def timing(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_size_t()
        check(self._f("get_timing")(self.handle, C.byref(ms), C.byref(n)), f"{self._pre}_get_timing")
        return ms.value, n.value

    @property
    def step(self) -> int:
        it = C.c_size_t()
        check(self._f("current_step")(self.handle, C.byref(it)), f"{self._pre}_current_step")
        return it.value

    def traces(self, it0: int, count: int) -> np.ndarray:
        out = np.zeros((count, self.nrec), dtype=np.float64)
        check(self._f("get_traces")(self.handle, it0, count, _dp(out)), f"{self._pre}_get_traces")
        return out

    def field(self, name: str) -> np.ndarray:
        out = np.zeros(self.cfg.shape, dtype=np.float64)
        check(self._f("get_field")(self.handle, self.FIELDS[name], _dp(out)), f"{self._pre}_get_field")
        return out

    def reset(self) -> None:
        check(self._f("reset")(self.handle), f"{self._pre}_reset")

    def device_bytes(self) -> int:
        b = C.c_size_t()
        check(self._f("device_bytes")(self.handle, C.byref(b)), f"{self._pre}_device_bytes")
        return b.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(self._f("destroy")(self.handle), f"{self._pre}_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ViscoelasticSH(_Pstd):
    """Viscoelastic SH waves in a 2D VTI medium (ve_vti_sh_pml): pseudo-
    spectral derivatives, CPML, fourth-order Runge-Kutta."""

    _pre = "sm_sh"
    FIELDS = {"vy": L.SM_VY, "txy": L.SM_TXY, "tyz": L.SM_TYZ}

    def __init__(self, ctx: Context, cfg: ShConfig, mu, rho, gamma, Q, f):
        self.ctx, self.cfg = ctx, cfg
        n = int(np.prod(cfg.shape))
        self._m = [_f64(a).reshape(-1) for a in (mu, rho, gamma, Q, f)]
        if any(a.size != n for a in self._m):
            raise ValueError(f"model must hold {n} padded cells")
        self._p = cfg.params()
        h = C.c_void_p()
        ctx.adopt(self)
        check(lib.sm_sh_create(ctx.handle, C.byref(self._p), *(_dp(a) for a in self._m), C.byref(h)),
              "sm_sh_create")
        self.handle = h
        self.nrec = 0


@dataclass
class VeConfig:
    """ve_vti_pml grid: nx by nz interior nodes and np CPML layers on every
    side (none when abs is False), padded (nx + 2np) by (nz + 2np); L
    quasi-dilatational relaxation mechanisms."""
    nx: int
    nz: int
    np: int
    dx: float
    dz: float
    dt: float
    f0: float
    L: int = 1
    vmax: float = 5000.0
    pml_order: float = 2.0
    Rc: float = 1e-5
    kappa_max: float = 1.0
    alpha_max: int = 1
    abs: bool = True
    storage: str = "f64"
    hipfft: bool = False      # every derivative through hipFFT (SM_PSTD_HIPFFT)

    def params(self) -> L.VeParams:
        return L.VeParams(self.nx, self.nz, self.np, self.dx, self.dz, self.dt, self.vmax,
                          self.pml_order, self.Rc, self.kappa_max, self.alpha_max, self.f0,
                          1 if self.abs else 0, self.L, _STORAGE[self.storage],
                          L.SM_PSTD_HIPFFT if self.hipfft else 0)

    @property
    def shape(self) -> tuple[int, int]:
        np_ = self.np if self.abs else 0
        return (self.nx + 2 * np_, self.nz + 2 * np_)

    def bytes_needed(self, nrec: int = 0, max_steps: int = 0) -> int:
        p = self.params()
        return int(lib.sm_ve_bytes_needed(C.byref(p), nrec, max_steps))


def ve_moduli(L_: int, K, mu, rho, eps, delta, Q1, f1, Q2, f2) -> dict:
    """c11, c13, c33, c55 and K0 (MPa) as ve_vti_pml writes them
    (ve_vti_pml.c:695-699), from read_model_ve's arrays; host only."""
    K = _f64(K)
    nx2, nz2 = K.shape
    arrs = [_f64(a).reshape(-1) for a in (K, mu, rho, eps, delta, Q1, f1, Q2, f2)]
    out = {k: np.zeros((nx2, nz2)) for k in ("c11", "c13", "c33", "c55", "K0")}
    check(lib.sm_ve_moduli(nx2, nz2, L_, *(_dp(a) for a in arrs), *(_dp(o) for o in out.values())),
          "sm_ve_moduli")
    return out


class ViscoelasticPSV(_Pstd):
    """Viscoelastic P-SV waves in a 2D VTI medium (ve_vti_pml): pseudo-
    spectral derivatives, CPML, fourth-order Runge-Kutta, L + 1 relaxation
    mechanisms."""

    _pre = "sm_ve"
    FIELDS = {"vx": L.SM_VX, "vz": L.SM_VZ, "txx": L.SM_TXX, "tzz": L.SM_TZZ, "txz": L.SM_TXZ}

    def __init__(self, ctx: Context, cfg: VeConfig, K, mu, rho, eps, delta, Q1, f1, Q2, f2):
        self.ctx, self.cfg = ctx, cfg
        n = int(np.prod(cfg.shape))
        self._m = [_f64(a).reshape(-1) for a in (K, mu, rho, eps, delta, Q1, f1, Q2, f2)]
        want = [n, n, n, n, n, cfg.L * n, cfg.L * n, n, n]
        if any(a.size != w for a, w in zip(self._m, want)):
            raise ValueError(f"model must hold {n} padded cells ({cfg.L} arrays for Q1 and f1)")
        self._p = cfg.params()
        h = C.c_void_p()
        ctx.adopt(self)
        check(lib.sm_ve_create(ctx.handle, C.byref(self._p), *(_dp(a) for a in self._m), C.byref(h)),
              "sm_ve_create")
        self.handle = h
        self.nrec = 0


@dataclass
class PveConfig(ShConfig):
    """pve_iso_pml grid: nx by nz interior nodes and np CPML layers on every
    side (none when abs is False), padded (nx + 2np) by (nz + 2np)."""

    pml_median: bool = False  # 3 x 3 median smoothing of the layers' material (SM_PVE_PML_MEDIAN)

    def params(self) -> L.PveParams:
        p = super().params()
        if self.pml_median:
            p.flags |= L.SM_PVE_PML_MEDIAN
        return p

    def bytes_needed(self, nrec: int = 0, max_steps: int = 0) -> int:
        p = self.params()
        return int(lib.sm_pve_bytes_needed(C.byref(p), nrec, max_steps))


class Poroviscoelastic(_Pstd):
    """Poroviscoelastic waves in a 2D isotropic medium (pve_iso_pml): Biot
    solid and fluid motion, one relaxation mechanism for the fluid pressure,
    pseudo-spectral derivatives, CPML, fourth-order Runge-Kutta with the
    viscous coupling stepped analytically.  Model arrays as read_model
    leaves them: K_m, K_s, K_f, mu (GPa), phi, rho_s, rho_f (kg/m^3), T,
    eta (cP), kappa (mD), Q, fq (Hz)."""

    _pre = "sm_pve"
    FIELDS = {"vx": L.SM_VX, "vz": L.SM_VZ, "qx": L.SM_QX, "qz": L.SM_QZ, "txx": L.SM_TXX,
              "tzz": L.SM_TZZ, "txz": L.SM_TXZ, "p": L.SM_P, "div": L.SM_DIV, "curl": L.SM_CURL}

    def __init__(self, ctx: Context, cfg: PveConfig, K_m, K_s, K_f, phi, mu, rho_s, rho_f, T, eta, kappa,
                 Q, fq):
        self.ctx, self.cfg = ctx, cfg
        n = int(np.prod(cfg.shape))
        self._m = [_f64(a).reshape(-1) for a in (K_m, K_s, K_f, phi, mu, rho_s, rho_f, T, eta, kappa, Q, fq)]
        if any(a.size != n for a in self._m):
            raise ValueError(f"model must hold {n} padded cells")
        self._p = cfg.params()
        h = C.c_void_p()
        ctx.adopt(self)
        check(lib.sm_pve_create(ctx.handle, C.byref(self._p), *(_dp(a) for a in self._m), C.byref(h)),
              "sm_pve_create")
        self.handle = h
        self.nrec = 0
        self.energy_steps = 0

    def set_energy(self, i1: int, i2: int, j1: int, j2: int, rho11, rho12, rho22, max_steps: int) -> None:
        """Record the kinetic energy of padded cells [i1, i2] x [j1, j2] after
        every step; rho11, rho12, rho22 hold one value per region cell."""
        ni, nj = i2 - i1 + 1, j2 - j1 + 1
        a = [_f64(r).reshape(-1) for r in (rho11, rho12, rho22)]
        if any(r.size != ni * nj for r in a):
            raise ValueError(f"energy coefficients must hold {ni * nj} values")
        check(lib.sm_pve_set_energy(self.handle, i1, i2, j1, j2, *(_dp(r) for r in a), max_steps),
              "sm_pve_set_energy")
        self.energy_steps = max_steps

    def set_energy_axes(self, i1: int, i2: int, j1: int, j2: int, ax, bx, cx, az, bz, cz, max_steps: int) -> None:
        """As set_energy with separate coefficients along x and z: the sum of
        0.5 ax vx^2 + bx vx qx + 0.5 cx qx^2 and the z counterparts."""
        ni, nj = i2 - i1 + 1, j2 - j1 + 1
        a = [_f64(r).reshape(-1) for r in (ax, bx, cx, az, bz, cz)]
        if any(r.size != ni * nj for r in a):
            raise ValueError(f"energy coefficients must hold {ni * nj} values")
        check(lib.sm_pve_set_energy_axes(self.handle, i1, i2, j1, j2, *(_dp(r) for r in a), max_steps),
              "sm_pve_set_energy_axes")
        self.energy_steps = max_steps

    def get_state(self) -> bytes:
        """The checkpoint blob the run continues from (sm_pve_get_state)."""
        n = C.c_size_t()
        check(lib.sm_pve_state_bytes(self.handle, C.byref(n)), "sm_pve_state_bytes")
        buf = C.create_string_buffer(n.value)
        check(lib.sm_pve_get_state(self.handle, buf, n.value), "sm_pve_get_state")
        return buf.raw

    def set_state(self, blob: bytes) -> None:
        check(lib.sm_pve_set_state(self.handle, blob, len(blob)), "sm_pve_set_state")

    def energy(self, it0: int, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=np.float64)
        check(lib.sm_pve_get_energy(self.handle, it0, count, _dp(out)), "sm_pve_get_energy")
        return out


class PoroviscoelasticVTI(Poroviscoelastic):
    """Poroviscoelastic waves in a 2D VTI Biot medium (pve_vti_pml): the
    isotropic solver's time loop with c11, c13, c33 from Thomsen's epsilon
    and delta, tortuosity and permeability along x (T1, kappa1) and z (T3,
    kappa3).  Model arrays as read_modelVTI leaves them."""

    def __init__(self, ctx: Context, cfg: PveConfig, K_m, K_s, K_f, phi, mu, rho_s, rho_f, T1, T3, eta, kappa1,
                 kappa3, Q, fq, epsilon, delta):
        self.ctx, self.cfg = ctx, cfg
        n = int(np.prod(cfg.shape))
        self._m = [_f64(a).reshape(-1) for a in (K_m, K_s, K_f, phi, mu, rho_s, rho_f, T1, T3, eta, kappa1, kappa3,
                                                  Q, fq, epsilon, delta)]
        if any(a.size != n for a in self._m):
            raise ValueError(f"model must hold {n} padded cells")
        self._p = cfg.params()
        h = C.c_void_p()
        ctx.adopt(self)
        check(lib.sm_pve_create_vti(ctx.handle, C.byref(self._p), *(_dp(a) for a in self._m), C.byref(h)),
              "sm_pve_create_vti")
        self.handle = h
        self.nrec = 0
        self.energy_steps = 0
