"""GPU tests of the drop-in programs bin/e_iso and bin/a_iso_3d.

Each program runs a small deck written in the reference's formats (.par,
model, source and record files) with FP64 storage (SEISMOD_STORAGE=f64),
and its output tree is compared with the FP64 oracle run on the same
inputs:
  - trace files trc/<base>/<base>_<n>.trc must be text-identical to the
    oracle's samples printed in the reference's format ("%le    %le", t and
    1e3 x value, io_utils.c:1744-1768), at the steps the record predicate
    (e_iso.c:643-658) selects;
  - netCDF snapshots ssh/<base>/... must hold the oracle's field at the
    snapshot step in mm/s, converted as the reference does: rounded to FP32,
    then times 1e3 in FP32 (io_utils.c:1964, 2142).
The FP32 default storage is checked against the same oracle within the
FP32 tolerance of the other tests.
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest
from tolerances import FP32_REL_L2, note

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "seismod_amd" / "bin"
VX, VY, VZ, P = 0, 1, 2, 8


def run(prog, cwd, storage, args=("-p", "run.par")):
    env = dict(os.environ, SEISMOD_STORAGE=storage)
    r = subprocess.run([str(BIN / prog), "-v", *args], cwd=cwd, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "End of computation" in r.stdout
    return r


def trace_lines(values, dt, t0, rdt):
    from seismod_amd.solver import record_due
    return [f"{dt * (it + 1):e}    {1e3 * v:e}" for it, v in enumerate(values)
            if record_due(it, dt, t0, rdt)]


def read_trc(path):
    return Path(path).read_text().splitlines()


def nc_field(path):
    from scipy.io import netcdf_file
    with netcdf_file(str(path), "r", mmap=False) as f:
        return np.array(f.variables["z"].data, dtype=np.float64)


# ------------------------------------------------------------------ 2D --
def deck_2d(tmp, nx=40, nz=30, npml=5, h=5.0, dt_ms=0.4, time_s=0.06):
    rng = np.random.default_rng(11)
    vp = np.where(np.arange(nz)[None, :] < nz // 2, 2500.0, 3500.0) * np.ones((nx, 1))
    vs = vp / 1.75
    rho = np.where(np.arange(nz)[None, :] < nz // 2, 2000.0, 2400.0) * np.ones((nx, 1))
    vp = vp * (1 + 0.02 * rng.uniform(-1, 1, (nx, nz)))
    with open(tmp / "model.dat", "w") as f:
        f.write(f"{nx} {nz}\n{h} {h}\n0.0 0.0\nisotropic_elastic\n")
        for i in range(nx):
            for j in range(nz):
                f.write(f"{float(vp[i, j])!r} {float(vs[i, j])!r} {float(rho[i, j])!r}\n")
    (tmp / "src.dat").write_text("1\n1\nBulk\n1.0\n40.0\n100.0 50.0\n")
    recs = [("Vx", "Trace", "50.0 25.0", 0.0, dt_ms), ("Vz", "Trace", "150.0 75.0", 0.01, 2 * dt_ms),
            ("Vz", "Trace", "195.0 145.0", 0.0, dt_ms), ("Vz", "Snapshot", None, 0.03, 0.0)]
    with open(tmp / "rcv.dat", "w") as f:
        f.write(f"{len(recs)}\n")
        for comp, kind, xz, t0, rdt in recs:
            f.write(f"{comp}\n{kind}\n" + (f"{xz}\n" if xz else "") + f"{t0}\n{rdt}\n")
    (tmp / "run.par").write_text(
        "model.dat # model, file\nsrc.dat # source, file\nrcv.dat # output, records\n"
        f"d2 # basename, name\n{time_s} # time, s\n{dt_ms} # dt, ms\n1 # abs, on\n{npml} # nl, layers\n")
    src_idx = np.array([20 * nz + 10], dtype=np.uintp)
    rec = [(10 * nz + 5, VX, 0.0, dt_ms), (30 * nz + 15, VZ, 0.01, 2 * dt_ms),
           (39 * nz + 29, VZ, 0.0, dt_ms)]
    return dict(nx=nx, nz=nz, npml=npml, h=h, dt=dt_ms * 1e-3, nsteps=int(round(time_s / (dt_ms * 1e-3))),
                vp=vp, vs=vs, rho=rho, src_idx=src_idx, rec=rec, snap_t0=0.03)


def oracle_2d(d, nsteps):
    cfg = dict(nx=d["nx"], nz=d["nz"], npml=d["npml"], h=d["h"], dt=d["dt"], q1_compat=True)
    L, fct = O.source_functions([1.0], [40.0], d["dt"])
    ridx = np.array([r[0] for r in d["rec"]], dtype=np.uintp)
    rcomp = np.array([r[1] for r in d["rec"]], dtype=np.intc)
    tr, f, _ = O.e2d_run(cfg, d["vp"], d["vs"], d["rho"], np.array([6], dtype=np.intc),
                         d["src_idx"], L, fct, ridx, rcomp, nsteps)
    NX, NZ = d["nx"] + 2 * d["npml"], d["nz"] + 2 * d["npml"]
    vz = f[NX * NZ:2 * NX * NZ].reshape(NX, NZ)
    p = d["npml"]
    return tr, vz[p:p + d["nx"], p:p + d["nz"]]


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_e_iso_program(tmp_path, storage):
    from seismod_amd.solver import nsteps as sm_nsteps, record_due
    d = deck_2d(tmp_path)
    run("e_iso", tmp_path, storage)
    ns = sm_nsteps(0.06, d["dt"])
    tr, _ = oracle_2d(d, ns)
    for q, (_, _, t0, rdt) in enumerate(d["rec"]):
        got = read_trc(tmp_path / "trc" / "d2" / f"d2_{q + 1}.trc")
        want = trace_lines(tr[:, q], d["dt"], t0, rdt * 1e-3)
        assert len(got) == len(want) > 0
        if storage == "f64":
            assert got == want, q
        else:
            g = np.array([float(l.split()[1]) for l in got])
            w = np.array([float(l.split()[1]) for l in want])
            assert np.linalg.norm(g - w) <= FP32_REL_L2 * np.linalg.norm(w), q
    snaps = sorted((tmp_path / "ssh" / "d2").glob("d2_4_*.nc"))
    assert len(snaps) == 1
    it_snap = [it for it in range(ns) if record_due(it, d["dt"], d["snap_t0"], 0.0)]
    assert len(it_snap) == 1
    _, vz = oracle_2d(d, it_snap[0] + 1)
    z = nc_field(snaps[0])
    assert z.shape == (d["nz"], d["nx"])
    want = (vz.T.astype(np.float32) * np.float32(1e3)).astype(np.float64)
    if storage == "f64":
        assert np.array_equal(z, want)
    else:
        assert np.linalg.norm(z - want) <= FP32_REL_L2 * np.linalg.norm(want)


# ------------------------------------------------------------------ 3D --
def deck_3d(tmp, n=(20, 18, 16), np_=5, dx=10.0, dt_ms=0.8, time_s=0.08):
    nx, ny, nz = n
    rng = np.random.default_rng(12)
    vz = np.where(np.arange(nz) < nz // 2, 1800.0, 3000.0)
    vp = np.broadcast_to(vz, n) * (1 + 0.02 * rng.uniform(-1, 1, n))
    rho = 310.0 * vp ** 0.25
    with open(tmp / "model.dat", "w") as f:
        f.write(f"{nx} {ny} {nz}\n{dx} {dx} {dx}\n0.0 0.0 0.0\nisotropic_acoustic_3d\n")
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    f.write(f"{float(vp[i, j, k])!r} {float(rho[i, j, k])!r}\n")
    (tmp / "src.dat").write_text("1\n1\nBulk\n1.0\n30.0\n100.0 90.0 40.0\n")
    recs = [("Vz", "Trace", "50.0 40.0 30.0", 0.0, dt_ms), ("Vx", "Trace", "150.0 90.0 100.0", 0.008, dt_ms),
            ("P", "Trace", "100.0 120.0 70.0", 0.0, 2 * dt_ms), ("Vz", "Snapshot", None, 0.04, 0.0)]
    with open(tmp / "rcv.dat", "w") as f:
        f.write(f"{len(recs)}\n")
        for comp, kind, xyz, t0, rdt in recs:
            f.write(f"{comp}\n{kind}\n" + (f"{xyz}\n" if xyz else "") + f"{t0}\n{rdt}\n")
    (tmp / "run.par").write_text(
        "model.dat # model, file\nsrc.dat # source, file\nrcv.dat # output, records\n"
        f"d3 # basename, name\n{time_s} # time, s\n{dt_ms} # dt, ms\n1 # abs, on\n{np_} # nl, layers\n"
        "5000 # vmax, m/s\n2 # pml_order\n1e-5 # Rc\n1 # kappa_max\n1 # alpha_max\n")
    n2 = [m + 2 * np_ for m in n]
    cell = lambda i, j, k: ((i + np_) * n2[1] + j + np_) * n2[2] + k + np_
    rec = [(cell(5, 4, 3), VZ, 0.0, dt_ms), (cell(15, 9, 10), VX, 0.008, dt_ms),
           (cell(10, 12, 7), P, 0.0, 2 * dt_ms)]
    vpp = np.pad(vp, np_, mode="edge")
    rhop = np.pad(rho, np_, mode="edge")
    return dict(n=n, np=np_, dx=dx, dt=dt_ms * 1e-3, vp=vpp, rho=rhop,
                src_idx=np.array([cell(10, 9, 4)], dtype=np.uintp), rec=rec, snap_t0=0.04)


def oracle_3d(d, nsteps):
    nx, ny, nz = d["n"]
    cfg = dict(nx=nx, ny=ny, nz=nz, np=d["np"], dx=d["dx"], dy=d["dx"], dz=d["dx"], dt=d["dt"],
               f0=30.0, vmax=5000.0, pml_order=2.0, Rc=1e-5, kappa_max=1.0, alpha_max=1)
    L, fct = O.source_functions([1.0], [30.0], d["dt"])
    ridx = np.array([r[0] for r in d["rec"]], dtype=np.uintp)
    rcomp = np.array([r[1] for r in d["rec"]], dtype=np.intc)
    return O.a3d_run(cfg, d["vp"], d["rho"], d["src_idx"], L, fct, ridx, rcomp, nsteps)


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_a_iso_3d_program(tmp_path, storage):
    from seismod_amd.solver import nsteps as sm_nsteps, record_due
    d = deck_3d(tmp_path)
    run("a_iso_3d", tmp_path, storage)
    ns = sm_nsteps(0.08, d["dt"])
    tr, _ = oracle_3d(d, ns)
    for q, (_, _, t0, rdt) in enumerate(d["rec"]):
        got = read_trc(tmp_path / "trc" / "d3" / f"d3_{q + 1}.trc")
        want = trace_lines(tr[:, q], d["dt"], t0, rdt * 1e-3)
        assert len(got) == len(want) > 0
        if storage == "f64":
            assert got == want, q
        else:
            g = np.array([float(l.split()[1]) for l in got])
            w = np.array([float(l.split()[1]) for l in want])
            assert np.linalg.norm(g - w) <= FP32_REL_L2 * np.linalg.norm(w), q
    snaps = sorted((tmp_path / "ssh" / "d3").glob("d3_4_*.nc"))
    assert len(snaps) == 1
    it_snap = [it for it in range(ns) if record_due(it, d["dt"], d["snap_t0"], 0.0)]
    assert len(it_snap) == 1
    _, f = oracle_3d(d, it_snap[0] + 1)
    nx, ny, nz = d["n"]
    np_ = d["np"]
    n2 = [m + 2 * np_ for m in d["n"]]
    nnn = n2[0] * n2[1] * n2[2]
    vz = f[3 * nnn:4 * nnn].reshape(n2)
    iy = int(90.0 / d["dx"])          # (y_src - y0)/dy without +np (a_iso_3d.c:726)
    sl = vz[np_:np_ + nx, iy, np_:np_ + nz]
    z = nc_field(snaps[0])
    assert z.shape == (nz, nx)
    want = (sl.T.astype(np.float32) * np.float32(1e3)).astype(np.float64)
    if storage == "f64":
        assert np.array_equal(z, want)
    else:
        assert np.linalg.norm(z - want) <= FP32_REL_L2 * np.linalg.norm(want) + 1e-30


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_a_iso_3d_segy_samples(tmp_path, storage):
    """segy 1: the velocity traces go to segy/<base>.sgy instead of trc/
    (a_iso_3d.c:454-462, 698-741).  Sample round(t/rdt) of each trace holds
    the recorded value rounded to FP32 (fill_data_segy, io_utils.c:1770-1774)
    at the steps the record predicate selects; the P record is not a SEG-Y
    trace (io_utils.c:3269-3559 writes Vx, Vy, Vz, div, curl only).  The
    reference's sample 0 is never written and its last one lands past the
    end of the buffer: here they are 0 and dropped (documented deviation)."""
    import struct
    from seismod_amd.solver import nsteps as sm_nsteps, record_due
    d = deck_3d(tmp_path)
    with open(tmp_path / "run.par", "a") as f:
        f.write("1 # segy, on\n12 # shotpoint, number\n")
    run("a_iso_3d", tmp_path, storage)
    assert not (tmp_path / "trc" / "d3" / "d3_1.trc").exists()
    ns = sm_nsteps(0.08, d["dt"])
    tr, _ = oracle_3d(d, ns)
    b = (tmp_path / "segy" / "d3.sgy").read_bytes()
    pos = 3600
    ntr = 0
    for q, (_, comp, t0, rdt_ms) in enumerate(d["rec"]):
        if comp == P:
            continue
        rdt = rdt_ms * 1e-3
        nsr = int(round(0.08 / rdt))
        hdr = b[pos:pos + 240]
        assert struct.unpack_from("<i", hdr, 0)[0] == ntr + 1            # trace sequence number
        assert struct.unpack_from("<i", hdr, 16)[0] == 12                # shot point
        assert struct.unpack_from("<h", hdr, 114)[0] == nsr
        got = np.frombuffer(b, dtype="<f4", count=nsr, offset=pos + 240).astype(np.float64)
        want = np.zeros(nsr)
        for it in range(ns):
            if record_due(it, d["dt"], t0, rdt):
                s = int(round(d["dt"] * (it + 1) / rdt))
                if s < nsr:
                    want[s] = float(np.float32(tr[it, q]))
        assert np.abs(want).max() > 0, q
        if storage == "f64":
            assert np.array_equal(got, want), q
        else:
            assert np.linalg.norm(got - want) <= 4e-6 * np.linalg.norm(want), q
        pos += 240 + 4 * nsr
        ntr += 1
    assert ntr == 2 and pos == len(b)


# ------------------------------------------------------------ e_cyl_iso --
def deck_cyl(tmp, nx=24, nz=30, np_=5, h=2.0, dt_ms=0.25, time_s=0.03, mode=1, plstr=1,
             kurkjian=False):
    rng = np.random.default_rng(13)
    vp = np.where(np.arange(nz)[None, :] < nz // 2, 2500.0, 3200.0) * np.ones((nx, 1))
    vp = vp * (1 + 0.02 * rng.uniform(-1, 1, (nx, nz)))
    vs = vp / 1.8
    vs[:2, :] = 0.0                                   # fluid-filled borehole
    rho = np.where(np.arange(nx)[:, None] < 2, 1000.0, 2300.0) * np.ones((1, nz))
    with open(tmp / "model.dat", "w") as f:
        f.write(f"{nx} {nz}\n{h} {h}\n0.0 0.0\nisotropic_elastic_cyl\n")
        for j in range(nz):                           # r fastest (io_utils.c:974-1000)
            for i in range(nx):
                f.write(f"{float(vp[i, j])!r} {float(vs[i, j])!r} {float(rho[i, j])!r}\n")
    # the Kurkjian source on the axis in the fluid; the radial force off the
    # axis in the solid (a radial force on the axis of a fluid does not
    # couple to the stresses of the first cell)
    src_r = 0.0 if kurkjian else 6.0
    src = (f"1\n1\nKurkjian\n1e-6\n400.0\n{src_r} 30.0\n" if kurkjian
           else f"1\n1\nFr\n1.0\n400.0\n{src_r} 30.0\n")
    (tmp / "src.dat").write_text(src)
    recs = [("Vr", "Trace", "10.0 40.0", 0.0, dt_ms), ("Vz", "Trace", "20.0 10.0", 0.005, 2 * dt_ms),
            ("trr", "Trace", "6.0 30.0", 0.0, dt_ms), ("div", "Trace", "12.0 24.0", 0.0, dt_ms),
            ("curl", "Trace", "14.0 50.0", 0.0, dt_ms), ("Vtheta", "Trace", "8.0 36.0", 0.0, dt_ms),
            ("Vz", "Snapshot", None, 0.02, 0.0)]
    with open(tmp / "rcv.dat", "w") as f:
        f.write(f"{len(recs)}\n")
        for comp, kind, xz, t0, rdt in recs:
            f.write(f"{comp}\n{kind}\n" + (f"{xz}\n" if xz else "") + f"{t0}\n{rdt}\n")
    (tmp / "run.par").write_text(
        "model.dat # model, file\nsrc.dat # source, file\nrcv.dat # output, records\n"
        f"c1 # basename, name\n{time_s} # time, s\n{dt_ms} # dt, ms\n1 # abs, on\n{np_} # nl, layers\n"
        f"{plstr} # plstr, strips\n{mode} # azimuthal_mode, n\n")
    nx2, nz2 = nx + np_, nz + 2 * np_
    vpp = np.zeros((nx2, nz2)); vsp = np.zeros((nx2, nz2)); rhop = np.zeros((nx2, nz2))
    for a, b in ((vpp, vp), (vsp, vs), (rhop, rho)):
        a[:nx, np_:np_ + nz] = b
        a[:nx, :np_] = b[:, :1]
        a[:nx, np_ + nz:] = b[:, -1:]
        a[nx:, :] = a[nx - 1:nx, :]
    cell = lambda r, z: int(round(r / h)) * nz2 + int(round(z / h)) + np_
    comp = {"Vr": 13, "Vz": 2, "trr": 15, "div": 11, "curl": 12, "Vtheta": 14}
    rec = [(cell(*map(float, xz.split())), comp[c], t0, rdt) for c, k, xz, t0, rdt in recs if xz]
    return dict(nx=nx, nz=nz, np=np_, h=h, dt=dt_ms * 1e-3, vp=vpp, vs=vsp, rho=rhop, mode=mode,
                src=cell(src_r, 30.0), src_r=src_r, rec=rec, snap_t0=0.02, kurkjian=kurkjian,
                f0=400.0)


def oracle_cyl(d, nsteps):
    import math
    cfg = dict(nx=d["nx"], nz=d["nz"], np=d["np"], dx=d["h"], dz=d["h"], dt=d["dt"], n=d["mode"],
               alpha_max=math.pi * d["f0"])
    if d["kurkjian"]:
        c = d["src"]
        vp, vs, rho = (a.reshape(-1)[c] for a in (d["vp"], d["vs"], d["rho"]))
        fct = O.kurkjian_fct(rho * (vp * vp - 2. * vs * vs), d["h"], d["h"], d["src_r"], d["mode"],
                             1e-6, d["f0"], d["dt"])
        L, st = np.array([len(fct)], np.uintp), 12
    else:
        L, fct = O.source_functions([1.0], [d["f0"]], d["dt"])
        st = 9
    ridx = np.array([r[0] for r in d["rec"]], dtype=np.uintp)
    rcomp = np.array([r[1] for r in d["rec"]], dtype=np.intc)
    return O.cyl_run(cfg, d["vp"], d["vs"], d["rho"], np.array([st], np.intc),
                     np.array([d["src"]], np.uintp), L, fct, ridx, rcomp, nsteps)


# FP32 e_cyl_iso deck: the 400 Hz source is four cells per P wavelength, so
# the field is dominated by grid-scale oscillation and FP32 rounding in the
# differences shows more than on the well-sampled decks (Vz snapshot measured
# at 6.1e-6 relative L2 against the FP64 oracle); ten times that
CYL_DROPIN_FP32 = 6e-5


@pytest.mark.parametrize("storage,mode,plstr,kurk", [("f64", 1, 1, False), ("f64", 0, 0, True),
                                                     ("f32", 1, 1, False)])
def test_e_cyl_iso_program(tmp_path, storage, mode, plstr, kurk):
    from seismod_amd.solver import nsteps as sm_nsteps, record_due
    d = deck_cyl(tmp_path, mode=mode, plstr=plstr, kurkjian=kurk)
    run("e_cyl_iso", tmp_path, storage)
    ns = sm_nsteps(0.03, d["dt"])
    tr, _ = oracle_cyl(d, ns)
    for q, (_, _, t0, rdt) in enumerate(d["rec"]):
        got = read_trc(tmp_path / "trc" / "c1" / f"c1_{q + 1}.trc")
        want = trace_lines(tr[:, q], d["dt"], t0, rdt * 1e-3)
        assert len(got) == len(want) > 0
        if storage == "f64":
            assert got == want, q
        else:
            g = np.array([float(l.split()[1]) for l in got])
            w = np.array([float(l.split()[1]) for l in want])
            e = np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-300)
            note("e_cyl_iso_program_fp32_trace", record=q, rel_l2=e)
            assert e <= CYL_DROPIN_FP32, q
    snaps = sorted((tmp_path / "ssh" / "c1").glob("c1_7_*.nc"))
    assert len(snaps) == 1
    it_snap = [it for it in range(ns) if record_due(it, d["dt"], d["snap_t0"], 0.0)]
    _, f = oracle_cyl(d, it_snap[0] + 1)
    nx2, nz2 = d["nx"] + d["np"], d["nz"] + 2 * d["np"]
    vz = f[2 * nx2 * nz2:3 * nx2 * nz2].reshape(nx2, nz2)
    # with strips the whole (nx+np) x (nz+2np) grid; without, the reference
    # skips np rows in r as well (io_utils.c:1985-1991)
    sl = vz if plstr else vz[d["np"]:d["np"] + d["nx"], d["np"]:d["np"] + d["nz"]]
    z = nc_field(snaps[0])
    assert z.shape == (sl.shape[1], sl.shape[0])
    want = (sl.T.astype(np.float32) * np.float32(1e3)).astype(np.float64)
    if storage == "f64":
        assert np.array_equal(z, want)
    else:
        e = np.linalg.norm(z - want) / np.linalg.norm(want)
        note("e_cyl_iso_program_fp32_snapshot", rel_l2=e)
        assert e <= CYL_DROPIN_FP32


# --------------------------------------------------------- ve_vti_sh_pml --
def deck_sh(tmp, nx=48, nz=40, np_=8, h=5.0, dt_ms=0.5, time_s=0.06, segy=0, template=1):
    rng = np.random.default_rng(17)
    vs = np.where(np.arange(nz)[None, :] < nz // 2, 1400.0, 2000.0) * np.ones((nx, 1))
    vs = vs * (1 + 0.02 * rng.uniform(-1, 1, (nx, nz)))
    rho = np.where(np.arange(nz)[None, :] < nz // 2, 1900.0, 2300.0) * np.ones((nx, 1))
    gamma = np.where(np.arange(nz)[None, :] < nz // 2, 0.1, 0.2) * np.ones((nx, 1))
    Q = np.where(np.arange(nz)[None, :] < nz // 2, 30.0, 60.0) * np.ones((nx, 1))
    fr = np.full((nx, nz), 40.0)
    with open(tmp / "model.dat", "w") as f:
        f.write(f"{nx} {nz}\n{h} {h}\n0.0 0.0\nvti_sh_viscoelastic\n")
        for i in range(nx):
            for j in range(nz):
                f.write(" ".join(repr(float(a[i, j])) for a in (vs, rho, gamma, Q, fr)) + "\n")
    (tmp / "src.dat").write_text(f"2\n{template}\nSyz\n2.0\n60.0\n120.0 50.0\nSxy\n0.5\n60.0\n80.0 120.0\n")
    recs = [("Vy", "Trace", "60.0 25.0", 0.0, dt_ms), ("txy", "Trace", "150.0 75.0", 0.01, 2 * dt_ms),
            ("tyz", "Trace", "200.0 180.0", 0.0, dt_ms), ("Vy", "Snapshot", None, 0.03, 0.0)]
    if segy:
        recs = [r for r in recs if r[1] == "Trace"]
        recs = [(c, k, xz, 0.0, dt_ms) for c, k, xz, _, _ in recs]
    with open(tmp / "rcv.dat", "w") as f:
        f.write(f"{len(recs)}\n")
        for comp, kind, xz, t0, rdt in recs:
            f.write(f"{comp}\n{kind}\n" + (f"{xz}\n" if xz else "") + f"{t0}\n{rdt}\n")
    (tmp / "run.par").write_text(
        "model.dat # model, file\nsrc.dat # source, file\nrcv.dat # output, records\n"
        f"sh # basename, name\n{time_s} # time, s\n{dt_ms} # dt, ms\n1 # abs, on\n{np_} # nl, layers\n"
        f"{segy} # segy, format\n3 # kappa_max, PML\n")
    nx2, nz2 = nx + 2 * np_, nz + 2 * np_
    pad = lambda a: np.pad(a, np_, mode="edge")
    cell = lambda x, z: (int(round(x / h)) + np_) * nz2 + int(round(z / h)) + np_
    rec = [(cell(*map(float, xz.split())), {"Vy": 1, "txy": 9, "tyz": 10}[c], t0, rdt)
           for c, k, xz, t0, rdt in recs if xz]
    return dict(nx=nx, nz=nz, np=np_, h=h, dt=dt_ms * 1e-3, vs=vs, rho=rho, mu=pad(vs * vs * rho * 1e-9),
                rhop=pad(rho), gamma=pad(gamma), Q=pad(Q), fr=pad(fr), rec=rec, template=template,
                srcs=[(5, cell(120.0, 50.0), 2.0), (3, cell(80.0, 120.0), 0.5)], snap_t0=0.03)


def oracle_sh(d, nsteps):
    from oracle import oracle_sh as S
    import math
    nz2 = d["nz"] + 2 * d["np"]
    pts = []
    w = [0.5, 0.125, 0.25, 0.125, 0.25, 0.25, 0.125, 0.25, 0.125]
    di = [0, -1, 0, 1, -1, 1, -1, 0, 1]
    dj = [0, -1, -1, -1, 0, 0, 1, 1, 1]
    for typ, c, A in d["srcs"]:
        if d["template"] == 9:
            pts += [(typ, c + di[q] * nz2 + dj[q], w[q] * A) for q in range(9)]
        else:
            pts.append((typ, c, A))
    L, fct = O.source_functions([p[2] for p in pts], [60.0] * len(pts), 0.5 * d["dt"])
    fct = fct / (d["h"] * d["h"])
    off = np.concatenate([[0], np.cumsum(L)]).astype(np.int64)
    src = [(p[0], p[1], fct[off[k]:off[k + 1]]) for k, p in enumerate(pts)]
    cfg = dict(nx=d["nx"], nz=d["nz"], np=d["np"], dx=d["h"], dz=d["h"], dt=d["dt"], vmax=5000.0,
               kappa_max=3.0, alpha_max=math.pi * 60.0, iwipe=1)
    ridx = [r[0] for r in d["rec"]]
    rcomp = [r[1] for r in d["rec"]]
    return S.sh_run(cfg, d["mu"], d["rhop"], d["gamma"], d["Q"], d["fr"], src, ridx, rcomp, nsteps)


def _close_lines(got, want, q):
    """Trace lines agree in time text and in value to the printed precision
    (FFT rounding differs between the FP64 GPU path and the oracle)."""
    assert len(got) == len(want) > 0, q
    g = np.array([float(l.split()[1]) for l in got])
    w = np.array([float(l.split()[1]) for l in want])
    assert [l.split()[0] for l in got] == [l.split()[0] for l in want], q
    assert np.abs(g - w).max() <= 2e-6 * np.abs(w).max(), q


@pytest.mark.parametrize("template", [1, 9])
def test_ve_vti_sh_pml_program(tmp_path, template):
    from seismod_amd.solver import nsteps as sm_nsteps, record_due
    d = deck_sh(tmp_path, template=template)
    run("ve_vti_sh_pml", tmp_path, "f64")
    ns = sm_nsteps(0.06, d["dt"])
    tr, _ = oracle_sh(d, ns)
    for q, (_, _, t0, rdt) in enumerate(d["rec"]):
        got = read_trc(tmp_path / "trc" / "sh" / f"sh_{q + 1}.trc")
        _close_lines(got, trace_lines(tr[:, q], d["dt"], t0, rdt * 1e-3), q)
    snaps = sorted((tmp_path / "ssh" / "sh").glob("sh_4_*.nc"))
    assert len(snaps) == 1
    it_snap = [it for it in range(ns) if record_due(it, d["dt"], d["snap_t0"], 0.0)]
    _, W = oracle_sh(d, it_snap[0] + 1)
    p = d["np"]
    want = (W[2][p:p + d["nx"], p:p + d["nz"]].T.astype(np.float32) * np.float32(1e3)).astype(np.float64)
    z = nc_field(snaps[0])
    assert z.shape == want.shape
    assert np.abs(z - want).max() <= 1e-5 * np.abs(want).max()


def test_ve_vti_sh_pml_segy(tmp_path):
    """segy 1: the Vy trace goes to segy/<base>.sgy (the stress records are
    not SEG-Y traces, io_utils.c:3269-3559), sample round(t/rdt) holding the
    value rounded to FP32 (fill_data_segy, io_utils.c:1770-1774); equal to
    the oracle's to the FFT rounding of the two FP64 paths."""
    from seismod_amd.solver import nsteps as sm_nsteps, record_due
    d = deck_sh(tmp_path, segy=1)
    run("ve_vti_sh_pml", tmp_path, "f64")
    assert not (tmp_path / "trc" / "sh" / "sh_1.trc").exists()
    ns = sm_nsteps(0.06, d["dt"])
    tr, _ = oracle_sh(d, ns)
    b = (tmp_path / "segy" / "sh.sgy").read_bytes()
    pos, ntr = 3600, 0
    for q, (_, comp, t0, rdt_ms) in enumerate(d["rec"]):
        if comp != VY:
            continue
        rdt = rdt_ms * 1e-3
        nsr = int(round(0.06 / rdt))
        got = np.frombuffer(b, dtype="<f4", count=nsr, offset=pos + 240).astype(np.float64)
        want = np.zeros(nsr)
        for it in range(ns):
            if record_due(it, d["dt"], t0, rdt):
                k = int(round(d["dt"] * (it + 1) / rdt))
                if k < nsr:
                    want[k] = float(np.float32(tr[it, q]))
        assert np.abs(want).max() > 0, q
        assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max(), q
        pos += 240 + 4 * nsr
        ntr += 1
    assert ntr == 1 and pos == len(b)


def deck_ve(tmp, nx=44, nz=38, np_=8, h=5.0, dt_ms=0.25, time_s=0.05, template=1, L=2):
    rng = np.random.default_rng(23)
    top = np.arange(nz)[None, :] < nz // 2
    ones = np.ones((nx, 1))
    vp = np.where(top, 2400.0, 3200.0) * ones * (1 + 0.02 * rng.uniform(-1, 1, (nx, nz)))
    vs = np.where(top, 1300.0, 1800.0) * ones
    rho = np.where(top, 1900.0, 2300.0) * ones
    eps = np.where(top, 0.1, 0.2) * ones
    dlt = np.where(top, 0.05, 0.1) * ones
    Q1 = [np.where(top, 40.0 + 10 * m, 80.0) * ones for m in range(L)]
    f1 = [np.full((nx, nz), 20.0 + 15 * m) for m in range(L)]
    Q2 = np.where(top, 30.0, 60.0) * ones
    f2 = np.full((nx, nz), 40.0)
    with open(tmp / "model.dat", "w") as f:
        f.write(f"{nx} {nz}\n{h} {h}\n0.0 0.0\nvti_viscoelastic\n{L}\n")
        for i in range(nx):
            for j in range(nz):
                vals = [vp, vs, rho, eps, dlt] + [a for m in range(L) for a in (Q1[m], f1[m])] + [Q2, f2]
                f.write(" ".join(repr(float(a[i, j])) for a in vals) + "\n")
    (tmp / "src.dat").write_text(f"2\n{template}\nBulk\n2.0\n60.0\n110.0 60.0\nSxz\n0.5\n60.0\n70.0 120.0\n")
    recs = [("Vx", "Trace", "60.0 25.0", 0.0, dt_ms), ("Vz", "Trace", "150.0 75.0", 0.01, 2 * dt_ms),
            ("txx", "Trace", "190.0 160.0", 0.0, dt_ms), ("tzz", "Trace", "30.0 150.0", 0.0, dt_ms),
            ("txz", "Trace", "100.0 100.0", 0.0, dt_ms), ("Vz", "Snapshot", None, 0.025, 0.0)]
    with open(tmp / "rcv.dat", "w") as f:
        f.write(f"{len(recs)}\n")
        for comp, kind, xz, t0, rdt in recs:
            f.write(f"{comp}\n{kind}\n" + (f"{xz}\n" if xz else "") + f"{t0}\n{rdt}\n")
    (tmp / "run.par").write_text(
        "model.dat # model, file\nsrc.dat # source, file\nrcv.dat # output, records\n"
        f"ve # basename, name\n{time_s} # time, s\n{dt_ms} # dt, ms\n1 # abs, on\n{np_} # nl, layers\n"
        "3 # kappa_max, PML\n1 # plstr, strips\n")
    nx2, nz2 = nx + 2 * np_, nz + 2 * np_
    pad = lambda a: np.pad(a, np_, mode="edge")
    cell = lambda x, z: (int(round(x / h)) + np_) * nz2 + int(round(z / h)) + np_
    comp = {"Vx": 0, "Vz": 2, "txx": 5, "tzz": 6, "txz": 7}
    rec = [(cell(*map(float, xz.split())), comp[c], t0, rdt) for c, k, xz, t0, rdt in recs if xz]
    mu_pa = vs * vs * rho
    model = (pad(1e-9 * (vp * vp * rho - 4. / 3. * mu_pa)), pad(mu_pa * 1e-9), pad(rho), pad(eps), pad(dlt),
             np.stack([pad(q) for q in Q1]), np.stack([pad(q) for q in f1]), pad(Q2), pad(f2))
    return dict(nx=nx, nz=nz, np=np_, h=h, dt=dt_ms * 1e-3, model=model, rec=rec, template=template,
                srcs=[(6, cell(110.0, 60.0), 2.0), (4, cell(70.0, 120.0), 0.5)], snap_t0=0.025)


def oracle_ve(d, nsteps):
    from oracle import oracle_ve as V
    import math
    nz2 = d["nz"] + 2 * d["np"]
    pts = []
    w = [0.5, 0.125, 0.25, 0.125, 0.25, 0.25, 0.125, 0.25, 0.125]
    di = [0, -1, 0, 1, -1, 1, -1, 0, 1]
    dj = [0, -1, -1, -1, 0, 0, 1, 1, 1]
    for typ, c, A in d["srcs"]:
        if d["template"] == 9:
            pts += [(typ, c + di[q] * nz2 + dj[q], w[q] * A) for q in range(9)]
        else:
            pts.append((typ, c, A))
    L, fct = O.source_functions([p[2] for p in pts], [60.0] * len(pts), 0.5 * d["dt"])
    fct = fct / (d["h"] * d["h"])
    off = np.concatenate([[0], np.cumsum(L)]).astype(np.int64)
    src = [(p[0], p[1], fct[off[k]:off[k + 1]]) for k, p in enumerate(pts)]
    cfg = dict(nx=d["nx"], nz=d["nz"], np=d["np"], dx=d["h"], dz=d["h"], dt=d["dt"], vmax=5000.0,
               kappa_max=3.0, alpha_max=math.pi * 60.0, iwipe=1)
    mat = V.ve_material(*d["model"])
    tr, W = V.ve_run(cfg, mat, src, [r[0] for r in d["rec"]], [r[1] for r in d["rec"]], nsteps)
    return tr, W, mat


@pytest.mark.parametrize("template", [1, 9])
def test_ve_vti_pml_program(tmp_path, template):
    """The drop-in ve_vti_pml (FP64) against the oracle: every trace line,
    the Vz snapshot (with the absorbing strips) and the five moduli files
    of ve_vti_pml.c:695-699."""
    from seismod_amd.solver import nsteps as sm_nsteps, record_due
    d = deck_ve(tmp_path, template=template)
    run("ve_vti_pml", tmp_path, "f64")
    ns = sm_nsteps(0.05, d["dt"])
    tr, _, mat = oracle_ve(d, ns)
    for q, (_, _, t0, rdt) in enumerate(d["rec"]):
        got = read_trc(tmp_path / "trc" / "ve" / f"ve_{q + 1}.trc")
        _close_lines(got, trace_lines(tr[:, q], d["dt"], t0, rdt * 1e-3), q)
    snaps = sorted((tmp_path / "ssh" / "ve").glob("ve_6_*.nc"))
    assert len(snaps) == 1
    it_snap = [it for it in range(ns) if record_due(it, d["dt"], d["snap_t0"], 0.0)]
    _, W, _ = oracle_ve(d, it_snap[0] + 1)
    want = (W[4].T.astype(np.float32) * np.float32(1e3)).astype(np.float64)
    z = nc_field(snaps[0])
    assert z.shape == want.shape
    assert np.abs(z - want).max() <= 1e-5 * np.abs(want).max()
    for name in ("c11", "c13", "c33", "c55", "K0"):
        z = nc_field(tmp_path / f"ve_{name}.nc")
        want = mat[name].T.astype(np.float32).astype(np.float64)
        assert np.allclose(z, want, rtol=2.5e-7, atol=0), name   # one FP32 ulp


# ------------------------------------------------------- poroviscoelastic --
def deck_pve(tmp, nx=44, nz=38, np_=8, h=5.0, dt_ms=0.25, time_s=0.05, template=1, with_pml=False, vti=False,
             median=False):
    rng = np.random.default_rng(29)
    top = np.arange(nz)[None, :] < nz // 2
    ones = np.ones((nx, 1))
    two = lambda a, b: np.where(top, a, b) * ones
    vals = [two(8.0, 12.0) * (1 + 0.02 * rng.uniform(-1, 1, (nx, nz))), two(36.0, 38.0), two(2.25, 2.0),
            two(0.3, 0.2), two(5.0, 8.0), two(2650.0, 2700.0), two(1000.0, 1040.0), two(2.0, 2.5),
            two(1.0, 0.8), two(800.0, 300.0), two(40.0, 2e4), two(30.0, 50.0)]
    npad = lambda a: np.pad(a, np_, mode="edge")
    if vti:
        # file columns: ... T3 eta kappa3 Q f0 epsilon delta kappa3/kappa1 T3/T1
        vals = vals + [two(0.12, 0.2), two(0.05, 0.1), two(3.0, 1.5), two(1.4, 1.2)]
    if with_pml:
        file_vals = [npad(a) for a in vals]
        head = (f"{nx} {nz}\n{h} {h}\n0.0 0.0\nanisotropic__vti_pml\n" if vti else
                f"{nx + 2 * np_} {nz + 2 * np_}\n{h} {h}\n0.0 0.0\nisotropic_pml\n")
    else:
        file_vals = vals
        head = f"{nx} {nz}\n{h} {h}\n0.0 0.0\n" + ("anisotropic_vti\n" if vti else "isotropic\n")
    with open(tmp / "model.dat", "w") as f:
        f.write(head)
        for i in range(file_vals[0].shape[0]):
            for j in range(file_vals[0].shape[1]):
                f.write(" ".join(repr(float(a[i, j])) for a in file_vals) + "\n")
    (tmp / "src.dat").write_text(f"2\n{template}\nBulk\n2.0\n60.0\n110.0 60.0\nSf\n0.5\n60.0\n70.0 120.0\n")
    recs = [("Vx", "Trace", "60.0 25.0", 0.0, dt_ms), ("qz", "Trace", "150.0 75.0", 0.01, 2 * dt_ms),
            ("p", "Trace", "100.0 80.0", 0.0, dt_ms), ("txz", "Trace", "30.0 150.0", 0.0, dt_ms),
            ("div", "Trace", "120.0 90.0", 0.0, dt_ms), ("curl", "Trace", "80.0 100.0", 0.0, dt_ms),
            ("p", "Snapshot", None, 0.025, 0.0)]
    if vti:       # pve_vti_pml records no div or curl
        recs[4:6] = [("tzz", "Trace", "120.0 90.0", 0.0, dt_ms), ("qx", "Trace", "80.0 100.0", 0.0, dt_ms)]
    with open(tmp / "rcv.dat", "w") as f:
        f.write(f"{len(recs)}\n")
        for comp, kind, xz, t0, rdt in recs:
            f.write(f"{comp}\n{kind}\n" + (f"{xz}\n" if xz else "") + f"{t0}\n{rdt}\n")
    roi = (50.0, 75.0, 40.0, 60.0)
    (tmp / "run.par").write_text(
        "model.dat # model, file\nsrc.dat # source, file\nrcv.dat # output, records\n"
        f"pve # basename, name\n{time_s} # time, s\n{dt_ms} # dt, ms\n1 # abs, on\n{np_} # nl, layers\n"
        "3 # kappa_max, PML\n1 # plstr, strips\n2 # check_model, store\n1 # saveEnergy, E(t)\n"
        + ("1 # pml_median, smoothing\n" if median else "") +
        f"{roi[0]}/{roi[1]}/{roi[2]}/{roi[3]} # EnergyROI, region\n")
    nz2 = nz + 2 * np_
    cell = lambda x, z: (int(round(x / h)) + np_) * nz2 + int(round(z / h)) + np_
    comp = {"Vx": 0, "qz": 4, "p": 8, "txz": 7, "div": 11, "curl": 12, "tzz": 6, "qx": 3}
    rec = [(cell(*map(float, xz.split())), comp[c], t0, rdt) for c, k, xz, t0, rdt in recs if xz]
    ij = [(i, j) for i in range(int(round(roi[0] / h)), int(round(roi[1] / h)) + 1)
          for j in range(int(round(roi[2] / h)), int(round(roi[3] / h)) + 1)]
    model = [npad(a) for a in vals]
    if vti:       # the solver's order: ... T1 T3 eta kappa1 kappa3 Q f0 epsilon delta
        K_m, K_s, K_f, phi, mu, rs, rf, T3, eta, k3, Q, f0, e_, d_, rk, rt = model
        model = [K_m, K_s, K_f, phi, mu, rs, rf, T3 / rt, T3, eta, k3 / rk, k3, Q, f0, e_, d_]
    return dict(nx=nx, nz=nz, np=np_, h=h, dt=dt_ms * 1e-3, model=tuple(model), rec=rec, vti=vti, median=median,
                template=template, srcs=[(6, cell(110.0, 60.0), 2.0), (7, cell(70.0, 120.0), 0.5)], snap_t0=0.025,
                roi_cells=ij, raw=vals)


def oracle_pve(d, nsteps, extra_rec=()):
    from oracle import oracle_pve as B
    import math
    nz2 = d["nz"] + 2 * d["np"]
    pts = []
    w = [0.5, 0.125, 0.25, 0.125, 0.25, 0.25, 0.125, 0.25, 0.125]
    di = [0, -1, 0, 1, -1, 1, -1, 0, 1]
    dj = [0, -1, -1, -1, 0, 0, 1, 1, 1]
    for typ, c, A in d["srcs"]:
        if d["template"] == 9:
            pts += [(typ, c + di[q] * nz2 + dj[q], w[q] * A) for q in range(9)]
        else:
            pts.append((typ, c, A))
    L, fct = O.source_functions([p[2] for p in pts], [60.0] * len(pts), 0.5 * d["dt"])
    fct = fct / (d["h"] * d["h"])
    off = np.concatenate([[0], np.cumsum(L)]).astype(np.int64)
    src = [(p[0], p[1], fct[off[k]:off[k + 1]]) for k, p in enumerate(pts)]
    cfg = dict(nx=d["nx"], nz=d["nz"], np=d["np"], dx=d["h"], dz=d["h"], dt=d["dt"], vmax=5000.0,
               kappa_max=3.0, alpha_max=math.pi * 60.0, iwipe=1)
    if d["vti"]:
        mat = B.pve_vti_material(*d["model"], d["dt"])
    else:
        mat = B.pve_material(*d["model"], d["dt"], median_np=d["np"] if d["median"] else 0)
    rec = list(d["rec"]) + list(extra_rec)
    return B.pve_run(cfg, mat, src, [r[0] for r in rec], [r[1] for r in rec], nsteps)


@pytest.mark.parametrize("template,with_pml,median", [(1, False, False), (9, True, False), (1, False, True)])
def test_pve_iso_pml_program(tmp_path, template, with_pml, median):
    """The drop-in pve_iso_pml (FP64) against the oracle: every trace line
    (divergence and curl included), the p snapshot (with the absorbing
    strips), the kinetic energy of the region in E/pve_E.dat (eq. 7.145 of
    Carcione 2007 from the oracle's velocities at every region cell) and the
    check_model files at the source frequency."""
    from seismod_amd.solver import nsteps as sm_nsteps, record_due
    d = deck_pve(tmp_path, template=template, with_pml=with_pml, median=median)
    r = run("pve_iso_pml", tmp_path, "f64")
    assert "Max/min phase velocities" in r.stdout
    ns = sm_nsteps(0.05, d["dt"])
    nz2 = d["nz"] + 2 * d["np"]
    p = d["np"]
    roi = [(i + p) * nz2 + j + p for i, j in d["roi_cells"]]
    extra = [(c, k, 0.0, 0.0) for k in (0, 2, 3, 4) for c in roi]
    tr, _ = oracle_pve(d, ns, extra)
    for q, (_, _, t0, rdt) in enumerate(d["rec"]):
        got = read_trc(tmp_path / "trc" / "pve" / f"pve_{q + 1}.trc")
        _close_lines(got, trace_lines(tr[:, q], d["dt"], t0, rdt * 1e-3), q)
    # energy: rho11, rho12, rho22 of eq. 7.153-7.160 per region cell
    nr, nq = len(d["rec"]), len(roi)
    vx, vz, qx, qz = (tr[:, nr + k * nq:nr + (k + 1) * nq] for k in range(4))
    phi, rs, rf, T = (np.array([d["raw"][q][i, j] for i, j in d["roi_cells"]]) for q in (3, 5, 6, 7))
    r12 = -phi * rf * (T - 1)
    r11 = (1 - phi) * rs - r12
    r22 = phi * rf - r12
    E = (0.5 * r11 * (vx * vx + vz * vz) + r12 * (vx * qx + vz * qz) + 0.5 * r22 * (qx * qx + qz * qz)).sum(1)
    lines = (tmp_path / "E" / "pve_E.dat").read_text().splitlines()
    assert len(lines) == ns
    got = np.array([[float(x) for x in l.split()] for l in lines])
    assert np.allclose(got[:, 0], d["dt"] * (np.arange(ns) + 1), rtol=1e-6)
    assert np.abs(E).max() > 0
    assert np.abs(got[:, 1] - E).max() <= 1e-6 * np.abs(E).max()   # printed with 7 digits
    snaps = sorted((tmp_path / "ssh" / "pve").glob("pve_7_*.nc"))
    assert len(snaps) == 1
    it_snap = [it for it in range(ns) if record_due(it, d["dt"], d["snap_t0"], 0.0)]
    _, W = oracle_pve(d, it_snap[0] + 1)
    want = (W[3].T.astype(np.float32) * np.float32(1e3)).astype(np.float64)
    z = nc_field(snaps[0])
    assert z.shape == want.shape
    assert np.abs(z - want).max() <= 1e-5 * np.abs(want).max()
    import test_host_io as H
    vel, att = H._dispersion([a.reshape(-1) for a in d["model"]], 60.0)
    for k, name in enumerate(("vel_slowP", "vel_fastP", "vel_S", "att_slowP", "att_fastP", "att_S")):
        src = (vel if k < 3 else att)[k % 3].reshape(d["model"][0].shape)
        assert np.array_equal(nc_field(tmp_path / f"pve_{name}.nc"), src.T.astype(np.float32).astype(np.float64))



@pytest.mark.parametrize("prog", ["pve_iso_pml", "pve_vti_pml"])
def test_pve_programs_checkpoint_restart(tmp_path, prog):
    """chkpt_inc 60 saves chkpt/pve_60.dat, pve_120.dat and pve_180.dat
    (after iterations it > 0 with it % 60 == 0, pve_iso_pml.c:1406-1411);
    -c chkpt/pve_120.dat, run from a directory holding the same inputs,
    continues at iteration 121 (pve_iso_pml.c:525-531) and writes the same
    trace and energy files as the whole run, byte for byte; the snapshot due
    before the restart is not written again."""
    import shutil
    vti = prog == "pve_vti_pml"
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    deck_pve(a, vti=vti)
    deck_pve(b, vti=vti)
    for d in (a, b):
        with open(d / "run.par", "a") as f:
            f.write("60 # chkpt_inc, checkpoints\n")
    run(prog, a, "f64")
    assert sorted(x.name for x in (a / "chkpt").iterdir()) == ["pve_120.dat", "pve_180.dat", "pve_60.dat"]
    shutil.copy(a / "chkpt" / "pve_120.dat", b / "ck.dat")
    (b / "run.par").unlink()        # the parameters come from the checkpoint
    r = run(prog, b, "f64", ("-c", "ck.dat"))
    assert "Starting with checkpoint file ck.dat" in r.stdout
    names = sorted(x.name for x in (a / "trc" / "pve").iterdir())
    assert len(names) == 6 and names == sorted(x.name for x in (b / "trc" / "pve").iterdir())
    for n in names:
        assert (a / "trc" / "pve" / n).read_bytes() == (b / "trc" / "pve" / n).read_bytes(), n
    ea = (a / "E" / "pve_E.dat").read_bytes()
    from seismod_amd.solver import nsteps as sm_nsteps
    assert len(ea.splitlines()) == sm_nsteps(0.05, 0.25e-3) and ea == (b / "E" / "pve_E.dat").read_bytes()
    assert len(list((a / "ssh" / "pve").glob("pve_7_*.nc"))) == 1
    assert not (b / "ssh" / "pve").exists() or not list((b / "ssh" / "pve").glob("pve_7_*.nc"))
    assert sorted(x.name for x in (b / "chkpt").iterdir()) == ["pve_180.dat"]


@pytest.mark.parametrize("template,with_pml", [(1, False), (9, True)])
def test_pve_vti_pml_program(tmp_path, template, with_pml):
    """The drop-in pve_vti_pml (FP64) against the oracle: every trace line,
    the p snapshot and the kinetic energy of the region (eq. 7.171 of
    Carcione 2007 with the x and z tortuosities, pve_vti_pml.c:1112-1120 and
    1334-1345) and the check_model files along x and z (pve_vti_pml.c:
    931-1100)."""
    from seismod_amd.solver import nsteps as sm_nsteps, record_due
    d = deck_pve(tmp_path, template=template, with_pml=with_pml, vti=True)
    r = run("pve_vti_pml", tmp_path, "f64")
    assert "Max/min phase velocities in X" in r.stdout and "Max/min attenuation coefficient in Z" in r.stdout
    ns = sm_nsteps(0.05, d["dt"])
    nz2 = d["nz"] + 2 * d["np"]
    p = d["np"]
    roi = [(i + p) * nz2 + j + p for i, j in d["roi_cells"]]
    extra = [(c, k, 0.0, 0.0) for k in (0, 2, 3, 4) for c in roi]
    tr, _ = oracle_pve(d, ns, extra)
    for q, (_, _, t0, rdt) in enumerate(d["rec"]):
        got = read_trc(tmp_path / "trc" / "pve" / f"pve_{q + 1}.trc")
        _close_lines(got, trace_lines(tr[:, q], d["dt"], t0, rdt * 1e-3), q)
    nr, nq = len(d["rec"]), len(roi)
    vx, vz, qx, qz = (tr[:, nr + k * nq:nr + (k + 1) * nq] for k in range(4))
    phi, rs, rf, T3, rt = (np.array([d["raw"][q][i, j] for i, j in d["roi_cells"]]) for q in (3, 5, 6, 7, 15))
    T1 = T3 / rt
    f1, f2 = (1 - phi) * rs, phi * rf
    rx, rz = (T1 - 1) / (1 / phi - 1), (T3 - 1) / (1 / phi - 1)
    E = 0.5 * (f1 * (vx * vx + vz * vz) - rx * (vx - qx) ** 2 - rz * (vz - qz) ** 2 + f2 * (qx * qx + qz * qz)).sum(1)
    got = np.array([[float(x) for x in l.split()] for l in (tmp_path / "E" / "pve_E.dat").read_text().splitlines()])
    assert got.shape == (ns, 2) and np.abs(E).max() > 0
    assert np.abs(got[:, 1] - E).max() <= 1e-6 * np.abs(E).max()
    snaps = sorted((tmp_path / "ssh" / "pve").glob("pve_7_*.nc"))
    assert len(snaps) == 1
    it_snap = [it for it in range(ns) if record_due(it, d["dt"], d["snap_t0"], 0.0)]
    _, W = oracle_pve(d, it_snap[0] + 1)
    want = (W[3].T.astype(np.float32) * np.float32(1e3)).astype(np.float64)
    z = nc_field(snaps[0])
    assert z.shape == want.shape
    assert np.abs(z - want).max() <= 1e-5 * np.abs(want).max()
    import test_host_io as H
    vel, att = H._dispersion_vti(list(d["model"]), 60.0)
    for k, nm in enumerate(("qP1_x", "qP2_x", "qS_x", "qP1_z", "qP2_z", "qS_z")):
        for pre, arr in (("vel", vel), ("att", att)):
            z = nc_field(tmp_path / f"pve_{pre}_{nm}.nc")
            assert np.array_equal(z, arr[k].T.astype(np.float32).astype(np.float64)), (pre, nm)
