"""CPU tests of the drop-in programs' input/output layer (seismod_amd/host).

The formats are those of seismod's io_utils.c: the .par grammar
(process_args, io_utils.c:156-294), model/source/record readers
(io_utils.c:892-947, 1029-1153, 1432-1741), the trace text format
(1744-1768), netCDF classic snapshots (1822-2180) and SEG-Y (3269-3559).
Snapshots are read back with scipy's independent netCDF-3 reader.
"""
import ctypes as C
import os
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = C.CDLL(str(ROOT / "seismod_amd" / "lib" / "libsm_hostio.so"))
LIBC = C.CDLL(None)
BIN = ROOT / "seismod_amd" / "bin"


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("duration", C.c_double),
                ("modelfile", C.c_char * 80), ("sourcefile", C.c_char * 80),
                ("outputfile", C.c_char * 80), ("basename", C.c_char * 80),
                ("iwipe", C.c_int), ("plot_strips", C.c_int), ("segy", C.c_int),
                ("shotpt_no", C.c_int), ("simulate_cmp", C.c_int), ("np", C.c_size_t),
                ("vmax", C.c_double), ("pml_order", C.c_double), ("Rc", C.c_double),
                ("kappa_max", C.c_double), ("alpha_max", C.c_int), ("mode", C.c_int),
                ("check_model", C.c_int), ("chkpt_inc", C.c_int), ("pml_median", C.c_int),
                ("save_energy", C.c_int), ("roi", C.c_double * 4)]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_size_t), ("ny", C.c_size_t), ("nz", C.c_size_t),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double)]


class Source(C.Structure):
    _fields_ = [("type", C.c_int), ("A", C.c_double), ("f", C.c_double), ("x", C.c_double),
                ("y", C.c_double), ("z", C.c_double), ("idx", C.c_size_t)]


class Sources(C.Structure):
    _fields_ = [("nsrc", C.c_size_t), ("ntemplate", C.c_size_t), ("s", C.POINTER(Source))]


class Record(C.Structure):
    _fields_ = [("comp", C.c_int), ("type", C.c_int), ("x", C.c_double), ("y", C.c_double),
                ("z", C.c_double), ("t0", C.c_double), ("dt", C.c_double), ("idx", C.c_size_t),
                ("fid", C.c_void_p), ("data", C.POINTER(C.c_float)), ("nsamples", C.c_size_t)]


class Records(C.Structure):
    _fields_ = [("nrec", C.c_size_t), ("r", C.POINTER(Record)), ("basename", C.c_char * 80)]


def parse_par(tmp_path, text, three_d=False):
    par = tmp_path / "run.par"
    par.write_text(text)
    p = Params()
    LIB.sm_set_defaults(C.byref(p), 1 if three_d else 0)
    C.c_int.in_dll(LIBC, "optind").value = 1
    argv = (C.c_char_p * 4)(b"prog", b"-p", str(par).encode(), None)
    LIB.sm_process_args(3, argv, C.byref(p), None)
    return p


def test_par_grammar(tmp_path):
    p = parse_par(tmp_path, """
model.dat          # model, model file
src.dat            # source, source file
rcv.dat            # output, description of output records

shot7              # basename, common for all output files
0.9                # time, window in s
0.6                # dt, time step in ms
1                  # abs, absorbing boundary
12                 # nl, number of PML layers
3.5                # kappa_max, PML parameter
0                  # alpha_max, PML parameter
0.001              # Rc, reflection coefficient
1                  # segy,save traces in segy format
42                 # unknown_keyword, ignored
""")
    assert p.modelfile == b"model.dat" and p.sourcefile == b"src.dat" and p.outputfile == b"rcv.dat"
    assert p.basename == b"shot7"
    assert p.duration == 0.9 and p.dt == 0.6 * 0.001
    assert p.np == 12 and p.kappa_max == 3.5 and p.alpha_max == 0 and p.Rc == 0.001
    assert p.segy == 1


def test_par_defaults_and_abs0(tmp_path):
    p = parse_par(tmp_path, "m.dat # model,\n30 # nl, layers\n")
    # io_utils.c:63-98 defaults; abs 0 (the default) removes the layers (293)
    assert p.dt == 1e-4 and p.duration == 1.2 and p.basename == b"run01"
    assert p.np == 0 and p.vmax == 5000.0 and p.pml_order == 2.0 and p.Rc == 1e-5


def write_model3d(path, vp, rho, dx=10.0, origin=(0.0, 0.0, 0.0), column=False):
    nx, ny, nz = vp.shape
    with open(path, "w") as f:
        f.write(f"{nx} {ny} {nz}\n{dx} {dx} {dx}\n{origin[0]} {origin[1]} {origin[2]}\n"
                "isotropic_acoustic_3d\n")
        cols = [(0, 0)] if column else [(i, j) for i in range(nx) for j in range(ny)]
        for i, j in cols:
            for k in range(nz):
                f.write(f"{float(vp[i, j, k])!r} {float(rho[i, j, k])!r}\n")


@pytest.mark.parametrize("column", [False, True])
def test_read_model_a_pads_by_edge_replication(tmp_path, column):
    rng = np.random.default_rng(1)
    nx, ny, nz, np_ = 4, 3, 5, 2
    vp = rng.uniform(1500, 4500, (nx, ny, nz))
    rho = rng.uniform(1800, 2600, (nx, ny, nz))
    if column:
        vp[:] = vp[0:1, 0:1, :]
        rho[:] = rho[0:1, 0:1, :]
    fn = tmp_path / "m3.dat"
    write_model3d(fn, vp, rho, column=column)
    g = Grid()
    LIB.sm_read_grid3d(str(fn).encode(), C.byref(g))
    assert (g.nx, g.ny, g.nz, g.dx) == (nx, ny, nz, 10.0)
    n2 = [n + 2 * np_ for n in (nx, ny, nz)]
    out_vp = np.zeros(n2)
    out_rho = np.zeros(n2)
    LIB.sm_read_model_a(str(fn).encode(), C.byref(g), C.c_size_t(np_),
                        out_vp.ctypes.data_as(C.c_void_p), out_rho.ctypes.data_as(C.c_void_p))
    assert np.array_equal(out_vp, np.pad(vp, np_, mode="edge"))
    assert np.array_equal(out_rho, np.pad(rho, np_, mode="edge"))


def test_read_model_e_full_and_column(tmp_path):
    nx, nz = 5, 4
    rng = np.random.default_rng(2)
    vals = rng.uniform(1000, 3000, (nx, nz, 3))
    for column in (False, True):
        fn = tmp_path / f"m2_{column}.dat"
        with open(fn, "w") as f:
            f.write(f"{nx} {nz}\n5.0 5.0\n0.0 0.0\nisotropic_elastic\n")
            for i in range(1 if column else nx):
                for j in range(nz):
                    f.write(" ".join(repr(float(v)) for v in vals[i, j]) + "\n")
        g = Grid()
        LIB.sm_read_grid2d(str(fn).encode(), C.byref(g))
        a = [np.zeros((nx, nz)) for _ in range(3)]
        LIB.sm_read_model_e(str(fn).encode(), C.byref(g), *(x.ctypes.data_as(C.c_void_p) for x in a))
        expect = np.broadcast_to(vals[0:1], vals.shape) if column else vals
        for q in range(3):
            assert np.array_equal(a[q], expect[:, :, q])


def _read3d(fn, np_):
    g = Grid()
    LIB.sm_read_grid3d(str(fn).encode(), C.byref(g))
    n2 = [n + 2 * np_ for n in (g.nx, g.ny, g.nz)]
    vp, rho = np.zeros(n2), np.zeros(n2)
    LIB.sm_read_model_a(str(fn).encode(), C.byref(g), C.c_size_t(np_),
                        vp.ctypes.data_as(C.c_void_p), rho.ctypes.data_as(C.c_void_p))
    return vp, rho


def test_model_writer_round_trip_parallel_read(tmp_path):
    """sm_write_model_a with 17 digits, read back by the parallel reader
    (a 64x48x40 body of about 6 MB is split across threads): every value
    comes back bit for bit, padded as before."""
    rng = np.random.default_rng(5)
    nx, ny, nz, np_ = 64, 48, 40, 3
    vp = rng.uniform(1500.0, 4500.0, (nx, ny, nz)) * 10.0 ** rng.integers(-3, 4, (nx, ny, nz))
    rho = -rng.standard_normal((nx, ny, nz)) * 1e-300          # tiny and subnormal, signed
    fn = tmp_path / "w3.dat"
    dbl = C.POINTER(C.c_double)
    LIB.sm_write_model_a.argtypes = [C.c_char_p, C.c_size_t, C.c_size_t, C.c_size_t] + \
        [C.c_double] * 6 + [dbl, dbl, C.c_int]
    assert LIB.sm_write_model_a(str(fn).encode(), nx, ny, nz, 10.0, 10.0, 10.0, 0.0, 0.0, 0.0,
                                vp.ctypes.data_as(dbl), rho.ctypes.data_as(dbl), 17) == 0
    out_vp, out_rho = _read3d(fn, np_)
    assert np.array_equal(out_vp, np.pad(vp, np_, mode="edge"))
    assert np.array_equal(out_rho, np.pad(rho, np_, mode="edge"))
    # the 2D writer and reader
    vs = vp[:, 0, :] / 1.8
    fn2 = tmp_path / "w2.dat"
    LIB.sm_write_model_e.argtypes = [C.c_char_p, C.c_size_t, C.c_size_t] + [C.c_double] * 4 + \
        [dbl, dbl, dbl, C.c_int]
    a = [np.ascontiguousarray(x) for x in (vp[:, 0, :], vs, rho[:, 0, :])]
    assert LIB.sm_write_model_e(str(fn2).encode(), nx, nz, 5.0, 5.0, 0.0, 0.0,
                                *(x.ctypes.data_as(dbl) for x in a), 17) == 0
    g = Grid()
    LIB.sm_read_grid2d(str(fn2).encode(), C.byref(g))
    b = [np.zeros((nx, nz)) for _ in range(3)]
    LIB.sm_read_model_e(str(fn2).encode(), C.byref(g), *(x.ctypes.data_as(C.c_void_p) for x in b))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_model_reader_number_forms_match_correct_rounding(tmp_path):
    """Any whitespace between numbers, exponents, signs, leading dots, long
    mantissas, no final newline: each value equals Python's correctly
    rounded float() of the same text (as fscanf's %lf gives)."""
    rng = np.random.default_rng(6)
    nx, ny, nz = 3, 4, 5
    forms = ["{:.3f}", "{:.6e}", "{:+.2E}", "{:.25g}", "{:.0f}", "{:.17g}"]
    toks = []
    for q in range(2 * nx * ny * nz):
        v = float(rng.uniform(0.001, 9000.0))
        t = forms[q % len(forms)].format(v)
        if q % 7 == 3 and t.startswith("0."):
            t = t[1:]                                   # ".5" form
        toks.append(t)
    seps = [" ", "\t", "\r\n", "\n\n ", "   "]
    body = "".join(t + seps[q % len(seps)] for q, t in enumerate(toks)).rstrip()
    fn = tmp_path / "forms.dat"
    fn.write_text(f"{nx} {ny} {nz}\n10 10 10\n0 0 0\nisotropic_acoustic_3d\n" + body)
    vp, rho = _read3d(fn, 0)
    expect = np.array([float(t) for t in toks]).reshape(nx, ny, nz, 2)
    assert np.array_equal(vp, expect[..., 0]) and np.array_equal(rho, expect[..., 1])


def _read3d_subprocess(fn, np_):
    code = (f"import ctypes as C, numpy as np, sys; sys.path.insert(0, {str(ROOT / 'tests')!r});"
            f"import test_host_io as T; vp, rho = T._read3d({str(fn)!r}, {np_});"
            "np.save(sys.argv[1], np.stack([vp, rho]))")
    out = Path(str(fn) + ".npy")
    r = subprocess.run(["python3", "-c", code, str(out)], capture_output=True, text=True)
    return r.returncode, r.stderr, (np.load(out) if r.returncode == 0 else None)


@pytest.mark.parametrize("cut", ["plane0", "plane1", "short", "bad_token"])
def test_model_reader_column_and_error_cases(tmp_path, cut):
    """io_utils.c:1049-1072: the model is one z column (copied everywhere)
    when the numbers end right after column (i, 0) of an x plane; a body
    ending anywhere else, or with a token that is not a number, stops the
    program with the reference's message."""
    rng = np.random.default_rng(7)
    nx, ny, nz = 3, 2, 4
    vals = rng.uniform(1000.0, 3000.0, (nx, ny, nz, 2))
    toks = [repr(float(v)) for v in vals.reshape(-1)]
    keep = {"plane0": 2 * nz, "plane1": 2 * nz * (ny + 1), "short": 2 * nz + 3,
            "bad_token": len(toks)}[cut]
    body = toks[:keep]
    if cut == "bad_token":
        body[2 * nz * ny + 5] = "12x"
    fn = tmp_path / f"{cut}.dat"
    fn.write_text(f"{nx} {ny} {nz}\n10 10 10\n0 0 0\nisotropic_acoustic_3d\n" + "\n".join(body) + "\n")
    rc, err, arr = _read3d_subprocess(fn, 1)
    if cut in ("plane0", "plane1"):
        assert rc == 0, err
        col = np.broadcast_to(vals[0:1, 0:1], vals.shape)
        assert np.array_equal(arr[0], np.pad(col[..., 0], 1, mode="edge"))
        assert np.array_equal(arr[1], np.pad(col[..., 1], 1, mode="edge"))
    else:
        assert rc == 1 and "Error: something wrong with file" in err


def test_read_sources(tmp_path):
    g = Grid(nx=400, ny=1, nz=400, dx=5.0, dy=5.0, dz=5.0)
    fn = tmp_path / "src.dat"
    fn.write_text("2\n1\nBulk\n1.0\n25.0\n1000.0 250.0\nSz\n-2.0\n30.0\n1002.4 247.6\n")
    s = Sources()
    LIB.sm_read_source2d(str(fn).encode(), C.byref(g), C.byref(s))
    assert (s.nsrc, s.ntemplate) == (2, 1)
    assert (s.s[0].type, s.s[0].idx) == (6, 200 * 400 + 50)        # Bulk at node (200, 50)
    assert (s.s[1].type, s.s[1].A, s.s[1].idx) == (2, -2.0, 200 * 400 + 50)
    # 9-point template: weights 1/2 centre, 1/4 edges, 1/8 corners (io_utils.c:1494-1535)
    fn.write_text("1\n9\nBulk\n2.0\n25.0\n1000.0 250.0\n")
    LIB.sm_read_source2d(str(fn).encode(), C.byref(g), C.byref(s))
    A = [s.s[q].A for q in range(9)]
    assert A == [1.0, 0.25, 0.5, 0.25, 0.5, 0.5, 0.25, 0.5, 0.25]
    idx = sorted(s.s[q].idx for q in range(9))
    assert idx == sorted((200 + di) * 400 + 50 + dj for di in (-1, 0, 1) for dj in (-1, 0, 1))
    g3 = Grid(nx=256, ny=256, nz=256, dx=10.0, dy=10.0, dz=10.0)
    fn.write_text("1\n1\nBulk\n1.0\n15.0\n1280.0 1280.0 100.0\n")
    LIB.sm_read_source3d(str(fn).encode(), C.byref(g3), C.c_size_t(20), C.byref(s))
    n2 = 296
    assert s.s[0].idx == ((128 + 20) * n2 + 128 + 20) * n2 + 10 + 20


def test_read_records(tmp_path):
    g3 = Grid(nx=32, ny=30, nz=28, dx=10.0, dy=10.0, dz=10.0, x0=100.0, y0=-50.0, z0=0.0)
    fn = tmp_path / "rcv.dat"
    fn.write_text("3\nVz\nTrace\n154.9 -50.0 35.0\n0.0\n0.6\nP\nTrace\n100 0 0 0.1 1.2\n"
                  "Vx\nSnapshot\n0.05\n0.0\n")
    out = Records()
    LIB.sm_read_output3d(str(fn).encode(), C.byref(g3), C.c_size_t(3), C.byref(out))
    assert out.nrec == 3
    r0, r1, r2 = out.r[0], out.r[1], out.r[2]
    nx2, ny2, nz2 = 38, 36, 34
    # lround: x 5.49 -> 5, y 0, z 3.5 -> 4 (io_utils.c:1722-1725)
    assert (r0.comp, r0.type, r0.idx) == (2, 0, ((5 + 3) * ny2 + 3) * nz2 + 4 + 3)
    assert abs(r0.dt - 0.6e-3) < 1e-18 and r0.t0 == 0.0
    assert (r1.comp, r1.idx, r1.t0) == (8, ((0 + 3) * ny2 + 5 + 3) * nz2 + 3, 0.1)
    assert (r2.comp, r2.type, r2.t0, r2.dt) == (0, 1, 0.05, 0.0)


def test_snapshot_netcdf_roundtrip(tmp_path, monkeypatch):
    from scipy.io import netcdf_file
    monkeypatch.chdir(tmp_path)
    nx, nz, np_ = 7, 5, 2
    rng = np.random.default_rng(3)
    data = rng.standard_normal((nx + 2 * np_, nz + 2 * np_)) * 1e-3
    for plstr in (0, 1):
        LIB.sm_write_snapshot_nc(data.ctypes.data_as(C.c_void_p), C.c_double(0.125), C.c_size_t(nx),
                                 C.c_size_t(nz), C.c_size_t(np_), C.c_double(100.0),
                                 C.c_double(-20.0), C.c_double(10.0), C.c_double(10.0), b"shot",
                                 C.c_size_t(plstr), C.c_int(plstr))
        fn = tmp_path / "ssh" / "shot" / f"shot_{plstr + 1}_0.125000.nc"
        with netcdf_file(str(fn), "r", mmap=False) as f:
            sl = slice(None) if plstr else slice(np_, -np_)
            sub = data[sl, sl]
            z = f.variables["z"]
            assert z.dimensions == ("y", "x")
            assert np.array_equal(z[:], (sub.astype(np.float32) * np.float32(1e3)).T)
            x = f.variables["x"][:]
            y = f.variables["y"][:]
            off = np_ if plstr else 0
            assert np.allclose(x, 100.0 + (np.arange(len(x)) - off) * 10.0)
            assert np.allclose(y, -20.0 + (np.arange(len(y)) - off) * 10.0)
            assert f.Conventions == b"COARDS/CF-1.0\x00"[:len(f.Conventions)] or \
                f.Conventions.rstrip(b"\x00") == b"COARDS/CF-1.0"
            assert f.title.rstrip(b"\x00") == b"pve_iso"
            assert int(np.asarray(f.node_offset).ravel()[0]) == 0
            assert np.allclose(z.actual_range, [1e3 * sub.min(), 1e3 * sub.max()])
            assert float(np.asarray(z.scale_factor).ravel()[0]) == 1.0
            assert np.isnan(np.asarray(z._FillValue).ravel()[0])
            assert z.units.rstrip(b"\x00") == b"mm/s"


def test_segy_layout(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    p = Params()
    LIB.sm_set_defaults(C.byref(p), 1)
    p.basename = b"line1"
    p.duration = 0.01
    p.shotpt_no = 7
    ns = 10
    recs = (Record * 2)()
    bufs = []
    for q in range(2):
        recs[q].comp = 2
        recs[q].type = 0
        recs[q].x = 100.0 + 20 * q
        recs[q].z = 30.0
        recs[q].dt = 0.001
        arr = (C.c_float * (ns + 1))(*[float(q * 100 + s) for s in range(ns + 1)])
        bufs.append(arr)
        recs[q].data = C.cast(arr, C.POINTER(C.c_float))
        recs[q].nsamples = ns
    out = Records(nrec=2, r=C.cast(recs, C.POINTER(Record)), basename=b"line1")
    src = (Source * 1)()
    src[0].x = 60.0
    src[0].z = 5.0
    srcs = Sources(nsrc=1, ntemplate=1, s=C.cast(src, C.POINTER(Source)))
    LIB.sm_save_segy(C.byref(p), C.byref(out), C.byref(srcs))
    b = (tmp_path / "segy" / "line1.sgy").read_bytes()
    assert len(b) == 3600 + 2 * (240 + 4 * ns)
    assert b[0:1] == b"C" and b[80:81] == b"C" and b[1:80] == b" " * 79
    h = lambda off: struct.unpack_from("<h", b, off)[0]
    i = lambda off: struct.unpack_from("<i", b, off)[0]
    assert (i(3200), h(3212), h(3216), h(3220), h(3224), h(3226)) == (1, 2, 1000, 10, 5, 2)
    assert h(3500) == 1
    t1 = 3600 + 240 + 4 * ns
    assert (i(3600), i(3616), h(3628), i(3636), i(3640), i(3648)) == (1, 7, 12, 40, -30, 5)
    assert (i(t1), i(t1 + 36)) == (2, 60)
    assert (h(3600 + 114), h(3600 + 116)) == (ns, 1000)
    samples = np.frombuffer(b, dtype="<f4", count=ns, offset=3840)
    assert np.array_equal(samples, np.arange(ns, dtype=np.float32))


def run_bin(prog, *args, cwd=None):
    return subprocess.run([str(BIN / prog), *args], capture_output=True, text=True, cwd=cwd,
                          timeout=60)


@pytest.mark.parametrize("prog", ["e_iso", "a_iso_3d"])
def test_cli_errors(prog, tmp_path):
    assert run_bin(prog, "-h").returncode == 0
    r = run_bin(prog)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = run_bin(prog, "-p", str(tmp_path / "missing.par"))
    assert r.returncode == 1 and "cannot open parameter file" in r.stderr
    assert run_bin(prog, "-c", "chk.dat").returncode == 1


def test_seeded_model_file_reads_back_the_shot(tmp_path):
    """synth.write_model_a writes a seeded shot's interior model; the
    drop-in reader returns exactly the padded arrays the shot holds (edge
    replication into the CPML is the reader's, as io_utils.c:1089-1143)."""
    import sys
    sys.path.insert(0, str(ROOT))
    from seismod_amd import synth
    shot = synth.a3d_layered_shot(n=24, np_=6, seed=3, vmax=5000.0)
    fn = tmp_path / "seeded.dat"
    synth.write_model_a(fn, shot)
    vp, rho = _read3d(fn, shot.np)
    assert np.array_equal(vp, shot.vp) and np.array_equal(rho, shot.rho)


def _read_positions_subprocess(tmp_path, src_text, rcv_text, np_):
    """Read a 3D source and record file in a child process (the readers exit
    on a bad position); returns (status, stderr, source idx, record idx)."""
    (tmp_path / "s.dat").write_text(src_text)
    (tmp_path / "r.dat").write_text(rcv_text)
    code = (f"import ctypes as C, sys; sys.path.insert(0, {str(ROOT / 'tests')!r});"
            "import test_host_io as T;"
            "g = T.Grid(nx=32, ny=30, nz=28, dx=10.0, dy=10.0, dz=10.0);"
            "s = T.Sources(); o = T.Records();"
            f"T.LIB.sm_read_source3d(b{str(tmp_path / 's.dat')!r}, C.byref(g), C.c_size_t({np_}), C.byref(s));"
            f"T.LIB.sm_read_output3d(b{str(tmp_path / 'r.dat')!r}, C.byref(g), C.c_size_t({np_}), C.byref(o));"
            "print(s.s[0].idx, o.r[0].idx)")
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True)
    idx = tuple(int(v) for v in r.stdout.split()) if r.returncode == 0 else None
    return r.returncode, r.stderr, idx


def test_3d_positions_inside_the_cpml_are_accepted(tmp_path):
    """a_iso_3d indexes sources and records by round(...) + np in the padded
    grid (io_utils.c:1582-1585, 1722-1725), so positions inside the CPML
    layers are valid; only positions past the padded grid, where the
    reference would index out of bounds, are refused."""
    np_, ny2, nz2 = 3, 36, 34
    rc, err, idx = _read_positions_subprocess(
        tmp_path, "1\n1\nBulk\n1.0\n15.0\n-20.0 100.0 -30.0\n",
        "1\nVz\nTrace\n330.0 -10.0 290.0 0.0 1.0\n", np_)
    assert rc == 0, err
    # source: x -2 -> 1, y 10 -> 13, z -3 -> 0; record: x 33 -> 36, y -1 -> 2, z 29 -> 32
    assert idx == (((1 * ny2) + 13) * nz2 + 0, ((36 * ny2) + 2) * nz2 + 32)
    rc, err, _ = _read_positions_subprocess(
        tmp_path, "1\n1\nBulk\n1.0\n15.0\n-40.0 100.0 0.0\n",
        "1\nVz\nTrace\n0.0 0.0 0.0 0.0 1.0\n", np_)
    assert rc == 1 and "outside the grid" in err
    rc, err, _ = _read_positions_subprocess(
        tmp_path, "1\n1\nBulk\n1.0\n15.0\n0.0 0.0 0.0\n",
        "1\nVz\nTrace\n0.0 0.0 310.0 0.0 1.0\n", np_)
    assert rc == 1 and "outside the grid" in err


@pytest.mark.parametrize("column", [False, True])
def test_read_model_e_cyl_layout_and_padding(tmp_path, column):
    """read_model_e_cyl (io_utils.c:949-1027): "Vp Vs rho" with r fastest
    into the (nx+np) x (nz+2np) grid; one r profile fills every z; edges
    copied into the z layers and the outer r layer (none at the axis)."""
    rng = np.random.default_rng(5)
    nx, nz, np_ = 5, 4, 2
    m = rng.uniform(1000, 3000, (3, nx, nz))
    if column:
        m[:] = m[:, :, :1]
    fn = tmp_path / "c.dat"
    with open(fn, "w") as f:
        f.write(f"{nx} {nz}\n1.5 2.0\n0.0 0.0\nisotropic_elastic_cyl\n")
        for j in range(1 if column else nz):
            for i in range(nx):
                f.write(" ".join(repr(float(m[q, i, j])) for q in range(3)) + "\n")
    g = Grid()
    LIB.sm_read_grid2d(str(fn).encode(), C.byref(g))
    out = [np.zeros((nx + np_, nz + 2 * np_)) for _ in range(3)]
    LIB.sm_read_model_e_cyl(str(fn).encode(), C.byref(g), C.c_size_t(np_),
                            *(o.ctypes.data_as(C.c_void_p) for o in out))
    for q in range(3):
        want = np.pad(m[q], ((0, np_), (np_, np_)), mode="edge")
        assert np.array_equal(out[q], want), q


def test_read_source_and_records_cyl(tmp_path):
    """Cylindrical index rules: row round((r - r0)/dr) with no axis layer,
    column round((z - z0)/dz) + np (io_utils.c:1488-1491, 1661-1666)."""
    g = Grid(nx=20, ny=1, nz=30, dx=2.0, dy=2.0, dz=2.0)
    np_ = 4
    nz2 = 30 + 2 * np_
    (tmp_path / "s.dat").write_text("1\n1\nKurkjian\n1e-6\n8000.0\n0.0 30.9\n")
    s = Sources()
    LIB.sm_read_source_cyl(str(tmp_path / "s.dat").encode(), C.byref(g), C.c_size_t(np_), C.byref(s))
    assert (s.s[0].type, s.s[0].idx) == (12, 0 * nz2 + 15 + np_)
    (tmp_path / "r.dat").write_text("2\nVr\nTrace\n7.1 -3.9 0.0 0.1\ntthetaz\nTrace\n40.0 10.0 0.0 0.1\n")
    out = Records()
    LIB.sm_read_output_cyl(str(tmp_path / "r.dat").encode(), C.byref(g), C.c_size_t(np_), C.byref(out))
    assert (out.r[0].comp, out.r[0].idx) == (13, 4 * nz2 - 2 + np_)     # z in the top layer
    assert (out.r[1].comp, out.r[1].idx) == (19, 20 * nz2 + 5 + np_)    # r in the outer layer


@pytest.mark.parametrize("column", [False, True])
def test_read_model_ve_sh_layout_and_padding(tmp_path, column):
    """read_grid_params_ve_sh / read_model_ve_sh (io_utils.c:330-354,
    1316-1429): "Vs0 rho gamma Q f" with z fastest into the (nx+2np) x
    (nz+2np) grid, mu = Vs0^2 rho 1e-9; one column fills every x; edges
    copied into all four layers; the keyword is case-insensitive in the
    header check only."""
    rng = np.random.default_rng(6)
    nx, nz, np_ = 4, 5, 2
    m = np.stack([rng.uniform(800, 2500, (nx, nz)), rng.uniform(1500, 2700, (nx, nz)),
                  rng.uniform(0, 0.3, (nx, nz)), rng.uniform(10, 90, (nx, nz)), rng.uniform(5, 50, (nx, nz))])
    if column:
        m[:] = m[:, :1, :]
    fn = tmp_path / "sh.dat"
    with open(fn, "w") as f:
        f.write(f"{nx} {nz}\n2.0 2.5\n0.0 0.0\nvti_sh_viscoelastic\n")
        for i in range(1 if column else nx):
            for j in range(nz):
                f.write(" ".join(repr(float(m[q, i, j])) for q in range(5)) + "\n")
    g = Grid()
    LIB.sm_read_grid_ve_sh(str(fn).encode(), C.byref(g))
    assert (g.nx, g.nz, g.dx, g.dz) == (nx, nz, 2.0, 2.5)
    out = [np.zeros((nx + 2 * np_, nz + 2 * np_)) for _ in range(5)]
    LIB.sm_read_model_ve_sh(str(fn).encode(), C.byref(g), C.c_size_t(np_),
                            *(o.ctypes.data_as(C.c_void_p) for o in out))
    mu = m[0] * m[0] * m[1] * 1e-9
    for q, want in enumerate([mu, m[1], m[2], m[3], m[4]]):
        assert np.array_equal(out[q], np.pad(want, np_, mode="edge")), q


def test_read_source_and_records_padded_2d(tmp_path):
    """ve_vti_sh_pml index rules: round((x - x0)/dx) + np along both axes of
    the padded grid, for sources (1- and 9-point templates) and records
    (io_utils.c:1488-1535, 1661-1666)."""
    g = Grid(nx=20, ny=1, nz=30, dx=2.0, dy=2.0, dz=2.0)
    np_ = 4
    nz2 = 30 + 2 * np_
    (tmp_path / "s.dat").write_text("1\n9\nSyz\n2.0\n25.0\n10.2 21.1\n")
    s = Sources()
    LIB.sm_read_source2d_np(str(tmp_path / "s.dat").encode(), C.byref(g), C.c_size_t(np_), C.byref(s))
    assert s.nsrc == 1 and s.ntemplate == 9
    i0, j0 = 5 + np_, 11 + np_
    want = [(i0, j0), (i0 - 1, j0 - 1), (i0, j0 - 1), (i0 + 1, j0 - 1), (i0 - 1, j0), (i0 + 1, j0),
            (i0 - 1, j0 + 1), (i0, j0 + 1), (i0 + 1, j0 + 1)]
    wts = [0.5, 0.125, 0.25, 0.125, 0.25, 0.25, 0.125, 0.25, 0.125]
    for q in range(9):
        assert (s.s[q].type, s.s[q].idx) == (5, want[q][0] * nz2 + want[q][1])
        assert s.s[q].A == wts[q] * 2.0
    (tmp_path / "r.dat").write_text("2\nVy\nTrace\n7.1 -3.9 0.0 0.1\ntxy\nTrace\n40.0 10.0 0.0 0.1\n")
    out = Records()
    LIB.sm_read_output2d_np(str(tmp_path / "r.dat").encode(), C.byref(g), C.c_size_t(np_), C.byref(out))
    assert (out.r[0].comp, out.r[0].idx) == (1, (4 + np_) * nz2 - 2 + np_)
    assert (out.r[1].comp, out.r[1].idx) == (9, (20 + np_) * nz2 + 5 + np_)


@pytest.mark.parametrize("column", [False, True])
def test_read_model_ve_layout_and_padding(tmp_path, column):
    """read_grid_params_ve / read_model_ve (io_utils.c:356-382, 1155-1314):
    per node "Vp0 Vs0 rho epsilon delta", L pairs "Q1 f1", "Q2 f2", z
    fastest, into the (nx+2np) x (nz+2np) grid; mu = Vs0^2 rho 1e-9 and
    K = 1e-9 (Vp0^2 rho - 4/3 Vs0^2 rho) in GPa; one column fills every x;
    edges copied into all four layers."""
    rng = np.random.default_rng(8)
    nx, nz, np_, L = 4, 5, 2, 2
    nv = 7 + 2 * L
    m = np.stack([rng.uniform(2000, 4500, (nx, nz)), rng.uniform(900, 2200, (nx, nz)),
                  rng.uniform(1500, 2700, (nx, nz)), rng.uniform(0, 0.3, (nx, nz)),
                  rng.uniform(-0.1, 0.2, (nx, nz))] + [rng.uniform(5, 90, (nx, nz)) for _ in range(2 * L + 2)])
    if column:
        m[:] = m[:, :1, :]
    fn = tmp_path / "ve.dat"
    with open(fn, "w") as f:
        f.write(f"{nx} {nz}\n2.0 2.5\n0.0 0.0\nvti_viscoelastic\n{L}\n")
        for i in range(1 if column else nx):
            for j in range(nz):
                f.write(" ".join(repr(float(m[q, i, j])) for q in range(nv)) + "\n")
    g = Grid()
    Lr = C.c_int()
    LIB.sm_read_grid_ve(str(fn).encode(), C.byref(g), C.byref(Lr))
    assert (g.nx, g.nz, g.dx, g.dz, Lr.value) == (nx, nz, 2.0, 2.5, L)
    shp = (nx + 2 * np_, nz + 2 * np_)
    out = [np.zeros(shp) for _ in range(5)] + [np.zeros((L,) + shp), np.zeros((L,) + shp)] + \
        [np.zeros(shp), np.zeros(shp)]
    LIB.sm_read_model_ve(str(fn).encode(), C.byref(g), C.c_size_t(np_), C.c_int(L),
                         *(o.ctypes.data_as(C.c_void_p) for o in out))
    vp, vs, rho = m[0], m[1], m[2]
    mu_pa = vs * vs * rho
    want = [1e-9 * (vp * vp * rho - 4. / 3. * mu_pa), mu_pa * 1e-9, rho, m[3], m[4]]
    for q in range(5):
        assert np.array_equal(out[q], np.pad(want[q], np_, mode="edge")), q
    for k in range(L):
        assert np.array_equal(out[5][k], np.pad(m[5 + 2 * k], np_, mode="edge"))
        assert np.array_equal(out[6][k], np.pad(m[6 + 2 * k], np_, mode="edge"))
    assert np.array_equal(out[7], np.pad(m[5 + 2 * L], np_, mode="edge"))
    assert np.array_equal(out[8], np.pad(m[6 + 2 * L], np_, mode="edge"))


def test_field_netcdf(tmp_path, monkeypatch):
    """write_field_nc (io_utils.c:2183-2352): <base>_<field>.nc in the current
    directory, values unscaled, units of strlen(units) bytes."""
    from scipy.io import netcdf_file
    monkeypatch.chdir(tmp_path)
    nx, nz, np_ = 6, 5, 2
    data = np.random.default_rng(4).uniform(1e3, 3e4, (nx + 2 * np_, nz + 2 * np_))
    LIB.sm_write_field_nc(data.ctypes.data_as(C.c_void_p), b"c11", b"MPa", C.c_size_t(nx), C.c_size_t(nz),
                          C.c_size_t(np_), C.c_double(10.0), C.c_double(0.0), C.c_double(5.0), C.c_double(5.0),
                          b"shot", C.c_int(1))
    with netcdf_file(str(tmp_path / "shot_c11.nc"), "r", mmap=False) as f:
        z = f.variables["z"]
        assert np.array_equal(z[:], data.astype(np.float32).T)
        assert z.units == b"MPa"
        assert np.allclose(z.actual_range, [data.min(), data.max()])
        assert np.allclose(f.variables["x"][:], 10.0 + (np.arange(nx + 2 * np_) - np_) * 5.0)


def test_ve_moduli_against_oracle():
    """sm_ve_moduli (host arithmetic of libseismod_hip, no GPU call) against
    the oracle's material (ve_vti_pml.c:518-648)."""
    from oracle import oracle_ve as V
    from seismod_amd.solver import ve_moduli
    rng = np.random.default_rng(9)
    shp, L = (9, 7), 2
    vp = rng.uniform(2200, 4000, shp); vs = vp / rng.uniform(1.6, 2.1, shp); rho = rng.uniform(1800, 2600, shp)
    model = (1e-9 * (vp * vp * rho - 4. / 3. * vs * vs * rho), vs * vs * rho * 1e-9, rho,
             rng.uniform(0, 0.25, shp), rng.uniform(-0.05, 0.15, shp), rng.uniform(20, 90, (L,) + shp),
             rng.uniform(5, 60, (L,) + shp), rng.uniform(15, 70, shp), rng.uniform(5, 60, shp))
    got = ve_moduli(L, *model)
    m = V.ve_material(*model)
    for k in ("c11", "c13", "c33", "c55", "K0"):
        assert np.allclose(got[k], m[k], rtol=1e-14, atol=0), k


def test_par_poroviscoelastic_keywords(tmp_path):
    """check_model, chkpt_inc, pml_median, saveEnergy and EnergyROI
    (io_utils.c:257-280)."""
    p = parse_par(tmp_path, "2 # check_model, store\n500 # chkpt_inc,\n1 # pml_median,\n"
                            "1 # saveEnergy, E(t)\n10/250/5.5/180 # EnergyROI, region\n")
    assert (p.check_model, p.chkpt_inc, p.pml_median, p.save_energy) == (2, 500, 1, 1)
    assert list(p.roi) == [10.0, 250.0, 5.5, 180.0]


def _read_pve(fn, np_):
    g = Grid()
    wp = C.c_int()
    LIB.sm_read_grid_pve(str(fn).encode(), C.c_size_t(np_), C.byref(g), C.byref(wp))
    shp = (g.nx + 2 * np_, g.nz + 2 * np_)
    out = [np.zeros(shp) for _ in range(12)]
    arr = (C.c_void_p * 12)(*(o.ctypes.data_as(C.c_void_p) for o in out))
    LIB.sm_read_model_pve(str(fn).encode(), C.byref(g), C.c_size_t(np_), wp, arr)
    return g, wp.value, np.stack(out)


def _read_pve_subprocess(fn, np_):
    code = (f"import numpy as np, sys; sys.path.insert(0, {str(ROOT / 'tests')!r});"
            f"import test_host_io as T; g, w, a = T._read_pve({str(fn)!r}, {np_});"
            "np.save(sys.argv[1], a)")
    out = Path(str(fn) + ".npy")
    r = subprocess.run(["python3", "-c", code, str(out)], capture_output=True, text=True)
    return r.returncode, r.stderr, (np.load(out) if r.returncode == 0 else None)


def _pve_values(rng, nx, nz):
    lo = [2, 30, 1.5, 0.1, 3, 2500, 900, 1.5, 0.5, 100, 10, 10]
    hi = [15, 40, 3.0, 0.35, 10, 2800, 1100, 3.0, 2.0, 2000, 100, 80]
    return np.stack([rng.uniform(a, b, (nx, nz)) for a, b in zip(lo, hi)])


@pytest.mark.parametrize("kind", ["full", "column", "with_pml"])
def test_read_model_pve_layout_and_padding(tmp_path, kind):
    """read_grid_params / read_model (io_utils.c:296-328, 403-625): 12 values
    per node, z fastest; "isotropic": one column fills every x and the
    nearest model node fills the layers; "isotropic_pml": the file holds the
    padded grid and its header counts the layers."""
    rng = np.random.default_rng(12)
    nx, nz, np_ = 4, 5, 2
    if kind == "with_pml":
        m = _pve_values(rng, nx + 2 * np_, nz + 2 * np_)
        head = f"{nx + 2 * np_} {nz + 2 * np_}\n2.0 2.5\n1.0 0.0\nIsotropic_PML\n"
    else:
        m = _pve_values(rng, nx, nz)
        if kind == "column":
            m[:] = m[:, :1, :]
        head = f"{nx} {nz}\n2.0 2.5\n1.0 0.0\nisotropic\n"
    fn = tmp_path / "pve.dat"
    with open(fn, "w") as f:
        f.write(head)
        for i in range(1 if kind == "column" else m.shape[1]):
            for j in range(m.shape[2]):
                f.write(" ".join(repr(float(m[q, i, j])) for q in range(12)) + "\n")
    g, wp, out = _read_pve(fn, np_)
    assert (g.nx, g.nz, g.dx, g.dz, g.x0, wp) == (nx, nz, 2.0, 2.5, 1.0, kind == "with_pml")
    for q in range(12):
        want = m[q] if kind == "with_pml" else np.pad(m[q], np_, mode="edge")
        assert np.array_equal(out[q], want), q


@pytest.mark.parametrize("cut", ["short", "bad_type", "pml_short", "pml_too_small"])
def test_read_model_pve_errors(tmp_path, cut):
    rng = np.random.default_rng(13)
    nx, nz, np_ = 4, 5, 2
    m = _pve_values(rng, nx + 2 * np_, nz + 2 * np_)
    toks = [repr(float(v)) for v in m.transpose(1, 2, 0).reshape(-1)]
    if cut == "short":
        head, body = f"{nx} {nz}\n2 2\n0 0\nisotropic\n", toks[:12 * (nz + 1)]
    elif cut == "bad_type":
        head, body = f"{nx} {nz}\n2 2\n0 0\nelastic\n", toks[:12 * nx * nz]
    elif cut == "pml_short":
        head, body = f"{nx + 2 * np_} {nz + 2 * np_}\n2 2\n0 0\nisotropic_pml\n", toks[:-3]
    else:
        head, body = f"{2 * np_} {nz}\n2 2\n0 0\nisotropic_pml\n", toks
    fn = tmp_path / f"{cut}.dat"
    fn.write_text(head + "\n".join(body) + "\n")
    rc, err, _ = _read_pve_subprocess(fn, np_)
    assert rc == 1
    assert {"short": "something wrong", "bad_type": "not isotropic", "pml_short": "something wrong",
            "pml_too_small": "not larger than its absorbing layers"}[cut] in err


def _dispersion(vals, f):
    n = vals[0].size
    vel = [np.zeros(n) for _ in range(3)]
    att = [np.zeros(n) for _ in range(3)]
    arr = (C.c_void_p * 12)(*(np.ascontiguousarray(v, np.float64).ctypes.data_as(C.c_void_p) for v in vals))
    LIB.sm_pve_dispersion(C.c_size_t(n), arr, C.c_double(f), (C.c_void_p * 3)(*(a.ctypes.data_as(C.c_void_p)
                                                                                for a in vel)),
                          (C.c_void_p * 3)(*(a.ctypes.data_as(C.c_void_p) for a in att)))
    return np.stack(vel), np.stack(att)


def test_pve_dispersion_limits():
    """The check_model velocities (pve_iso_pml.c:864-904) against closed
    forms independent of the dispersion relation's algebra:
      - lossless (eta = 0, Q >= 1e4): Biot's fast and slow P velocities, the
        square roots of the eigenvalues of R^-1 C (C the undrained stiffness
        [[E + alpha^2 M, alpha M], [alpha M, M]], R the density matrix
        [[rho, rho_f], [rho_f, T rho_f / phi]]), and S = sqrt(mu / (rho -
        phi rho_f / T)); no attenuation;
      - viscous fluid at low frequency (eta/kappa >> omega rho_f T / phi):
        the fast P velocity tends to Gassmann's sqrt((K_G + 4/3 mu) / rho)
        and S to sqrt(mu / rho) (the fluid moves with the frame);
      - a planted mistake (a sign of the coupling term) moves the lossless
        velocities by more than 1%."""
    rng = np.random.default_rng(21)
    n = 16
    vals = _pve_values(rng, 1, n).reshape(12, n)
    loss = vals.copy()
    loss[8] = 0.0
    loss[10] = 2e4
    vel, att = _dispersion(loss, 40.0)
    Km, Ks, Kf, phi, mu = loss[0] * 1e3, loss[1] * 1e3, loss[2] * 1e3, loss[3], loss[4] * 1e3
    rs, rf, T = loss[5] * 1e-6, loss[6] * 1e-6, loss[7]
    alpha = 1 - Km / Ks
    M = 1. / ((alpha - phi) / Ks + phi / Kf)
    E = Km + 4. / 3. * mu
    rho = (1 - phi) * rs + phi * rf
    for c in range(n):
        C_ = np.array([[E[c] + alpha[c] ** 2 * M[c], alpha[c] * M[c]], [alpha[c] * M[c], M[c]]])
        R = np.array([[rho[c], rf[c]], [rf[c], T[c] * rf[c] / phi[c]]])
        lam = np.sort(np.linalg.eigvals(np.linalg.solve(R, C_)).real)
        assert np.allclose(vel[:2, c], np.sqrt(lam), rtol=1e-12), c
        bad = np.array([[E[c] + alpha[c] ** 2 * M[c], -alpha[c] * M[c]], [-alpha[c] * M[c], M[c]]])
        lam_bad = np.sort(np.linalg.eigvals(np.linalg.solve(R, bad)).real)
        assert np.abs(np.sqrt(lam_bad) / vel[:2, c] - 1).max() > 1e-2
    assert np.allclose(vel[2], np.sqrt(mu / (rho - phi * rf / T)), rtol=1e-12)
    assert np.abs(att).max() < 1e-12
    visc = vals.copy()
    visc[10] = 2e4
    visc[8] = 1e3            # eta 1000 cP, kappa 100-2000 mD
    vel, _ = _dispersion(visc, 1e-3)
    Kg = Km + (1 - Km / Ks) ** 2 / (phi / Kf + (1 - phi) / Ks - Km / Ks ** 2)
    assert np.allclose(vel[1], np.sqrt((Kg + 4. / 3. * mu) / rho), rtol=1e-6)
    assert np.allclose(vel[2], np.sqrt(mu / rho), rtol=1e-6)


def test_process_args_with_checkpoints(tmp_path):
    """process_args (io_utils.c:156-171): -c names a checkpoint file and
    returns before any parameter file is read (-p is then ignored); without
    -c the .par is read and 0 returned."""
    par = tmp_path / "run.par"
    par.write_text("m.dat # model,\n7 # chkpt_inc,\n")
    ck = C.create_string_buffer(256)
    for argv_s, ret, name in [(["x", "-c", "chk.dat"], 1, b"chk.dat"),
                              (["x", "-p", str(tmp_path / "missing.par"), "-c", "a.dat"], 1, b"a.dat"),
                              (["x", "-p", str(par)], 0, b"")]:
        p = Params()
        LIB.sm_set_defaults(C.byref(p), 0)
        ck.value = b""
        argv = (C.c_char_p * (len(argv_s) + 1))(*[a.encode() for a in argv_s], None)
        C.c_int.in_dll(LIBC, "optind").value = 1   # getopt restarts at the first argument
        r = LIB.sm_process_args_ck(len(argv_s), argv, C.byref(p), None, ck, 256)
        assert (r, ck.value) == (ret, name)
        assert p.chkpt_inc == (7 if ret == 0 else 0)


def test_pve_programs_reject_foreign_checkpoints(tmp_path):
    """-c with a file that is not a checkpoint of the program stops with a
    message before a GPU is used; a pve_iso_pml checkpoint is refused by
    pve_vti_pml; both programs take -h."""
    for prog in ("pve_iso_pml", "pve_vti_pml"):
        assert run_bin(prog, "-h").returncode == 0
        r = run_bin(prog, "-c", "missing.dat", cwd=tmp_path)
        assert r.returncode == 1 and "cannot open checkpoint file" in r.stderr, r.stderr
    (tmp_path / "junk.dat").write_bytes(b"not a checkpoint at all" * 10)
    r = run_bin("pve_iso_pml", "-c", "junk.dat", cwd=tmp_path)
    assert r.returncode == 1 and "not a checkpoint of pve_iso_pml" in r.stderr, r.stderr
    # the header of an isotropic checkpoint: magic, medium 0, iteration,
    # parameters, state size 0
    blob = b"SMPVECK\0" + (0).to_bytes(4, "little") + (5).to_bytes(8, "little") + bytes(C.sizeof(Params)) + \
        (0).to_bytes(8, "little")
    (tmp_path / "iso.dat").write_bytes(blob)
    r = run_bin("pve_vti_pml", "-c", "iso.dat", cwd=tmp_path)
    assert r.returncode == 1 and "not a checkpoint of pve_vti_pml" in r.stderr, r.stderr


def _read_pve_vti(fn, np_):
    g = Grid()
    wp = C.c_int()
    LIB.sm_read_grid_pve_vti(str(fn).encode(), C.byref(g), C.byref(wp))
    shp = (g.nx + 2 * np_, g.nz + 2 * np_)
    out = [np.zeros(shp) for _ in range(16)]
    arr = (C.c_void_p * 16)(*(o.ctypes.data_as(C.c_void_p) for o in out))
    LIB.sm_read_model_pve_vti(str(fn).encode(), C.byref(g), C.c_size_t(np_), wp, arr)
    return g, wp.value, np.stack(out)


@pytest.mark.parametrize("kind", ["full", "column", "with_pml"])
def test_read_model_pve_vti_layout_and_padding(tmp_path, kind):
    """read_grid_params / read_modelVTI (io_utils.c:296-328, 627-888): 16
    values per node, kappa1 = kappa3 / column 15 and T1 = T3 / column 16;
    "anisotropic__vti_pml" holds the padded grid while its header still
    counts interior nodes (read_grid_params subtracts the layers only for
    "isotropic_pml")."""
    rng = np.random.default_rng(14)
    nx, nz, np_ = 4, 5, 2
    shp = (nx + 2 * np_, nz + 2 * np_) if kind == "with_pml" else (nx, nz)
    m = np.concatenate([_pve_values(rng, *shp), np.stack([rng.uniform(0, 0.3, shp), rng.uniform(-0.1, 0.2, shp),
                                                          rng.uniform(1, 5, shp), rng.uniform(0.5, 2, shp)])])
    if kind == "column":
        m[:] = m[:, :1, :]
    head = f"{nx} {nz}\n2.0 2.5\n1.0 0.0\n" + ("Anisotropic__VTI_pml\n" if kind == "with_pml" else "anisotropic_vti\n")
    fn = tmp_path / "pvevti.dat"
    # file columns: K_m K_s K_f phi mu rho_s rho_f T3 eta kappa3 Q f0 eps delta kappa3/kappa1 T3/T1
    with open(fn, "w") as f:
        f.write(head)
        for i in range(1 if kind == "column" else m.shape[1]):
            for j in range(m.shape[2]):
                f.write(" ".join(repr(float(m[q, i, j])) for q in range(16)) + "\n")
    g, wp, out = _read_pve_vti(fn, np_)
    assert (g.nx, g.nz, wp) == (nx, nz, kind == "with_pml")
    pad = (lambda a: a) if kind == "with_pml" else (lambda a: np.pad(a, np_, mode="edge"))
    order = {0: 0, 1: 1, 2: 2, 3: 3, 4: 4, 5: 5, 6: 6, 8: 7, 9: 8, 11: 9, 12: 10, 13: 11, 14: 12, 15: 13}
    for q_out, q_in in order.items():
        assert np.array_equal(out[q_out], pad(m[q_in])), q_out
    assert np.array_equal(out[10], pad(m[9] / m[14]))     # kappa1
    assert np.array_equal(out[7], pad(m[7] / m[15]))      # T1


def test_pve_vti_pml_usage():
    r = run_bin("pve_vti_pml", "-h")
    assert r.returncode == 0 and "anisotropic_vti" in r.stdout


def _vti_model(rng, shp, uniform=False):
    vals = list(_pve_values(rng, *shp))
    T3, K3 = rng.uniform(1.5, 3, shp), rng.uniform(100, 2000, shp)
    eps, dl = rng.uniform(0, 0.3, shp), rng.uniform(-0.05, 0.2, shp)
    # K_m K_s K_f phi mu rho_s rho_f T1 T3 eta kappa1 kappa3 Q f0 epsilon delta
    return [vals[0], vals[1], vals[2], vals[3], vals[4], vals[5], vals[6], vals[7], T3, vals[8], vals[9], K3,
            vals[10], vals[11], eps, dl]


def _dispersion_vti(v16, f):
    v16 = [np.ascontiguousarray(a, np.float64) for a in v16]
    nx2, nz2 = v16[0].shape
    vel = [np.zeros((nx2, nz2)) for _ in range(6)]
    att = [np.zeros((nx2, nz2)) for _ in range(6)]
    LIB.sm_pve_dispersion_vti(C.c_size_t(nx2), C.c_size_t(nz2),
                              (C.c_void_p * 16)(*(a.ctypes.data_as(C.c_void_p) for a in v16)), C.c_double(f),
                              (C.c_void_p * 6)(*(a.ctypes.data_as(C.c_void_p) for a in vel)),
                              (C.c_void_p * 6)(*(a.ctypes.data_as(C.c_void_p) for a in att)))
    return np.stack(vel), np.stack(att)


def _vti_check_inputs(v16):
    """The quantities pve_vti_pml hands to its velocity routines at check
    time: node material, eta/kappa averaged to the staggered x and z
    positions, the harmonic mean of mu at the cell corners."""
    from oracle import oracle_pve as B
    mat = B.pve_vti_material(*v16, 1e-4)
    phi, rs, rf = v16[3], v16[5] * 1e-6, v16[6] * 1e-6
    eta = v16[9] * B.ETA_SCALE
    nk1 = np.zeros_like(phi); nk1[:-1] = 0.5 * (eta[:-1] / v16[10][:-1] + eta[1:] / v16[10][1:]); nk1[-1] = nk1[-2]
    nk3 = np.zeros_like(phi)
    nk3[:, :-1] = 0.5 * (eta[:, :-1] / v16[11][:, :-1] + eta[:, 1:] / v16[11][:, 1:]); nk3[:, -1] = nk3[:, -2]
    return dict(mat=mat, rho=(1 - phi) * rs + phi * rf, rf=rf, m1=v16[7] * rf / phi, m3=v16[8] * rf / phi,
                nk1=nk1, nk3=nk3)


def test_pve_dispersion_vti_against_eigen_solutions_and_the_reference():
    """pve_vti_pml's check_model velocities (sm_pve_dispersion_vti) against
      - the exact plane-wave solutions along the symmetry axes, written
        independently: qP from the eigenvalues of R^-1 C with C = [[c11 +
        a1^2 M, a1 M], [a1 M, M]] and R = [[rho, rho_f], [rho_f, Y1]],
        Y1 = m1 - i (eta/kappa1)/omega (and the z counterparts), qS from
        c55 / (rho - rho_f^2 / Y) with the Y across the axis: 1e-10;
      - the reference's own VqP1, VqP2 and VqS (pve_vti_pml's general-
        direction closed forms, built from its sources into oracle/_ref) at
        (1, 0) and (0, 1), as complex slownesses: fast qP and qS within
        1e-5; the slow qP within 2%
        (the reference's closed form, a cubic solved by Cardano's formula,
        loses that much where the slow wave is diffusive: the eigen-solution
        above agrees with this implementation to 1e-13 there);
      - the isotropic limit (and Q >= 1e4: the VTI check leaves M
        unrelaxed): the velocities of sm_pve_dispersion."""
    rng = np.random.default_rng(3)
    shp = (5, 6)
    v16 = _vti_model(rng, shp)
    f = 40.0
    om = 2 * np.pi * f
    vel, att = _dispersion_vti(v16, f)
    q = _vti_check_inputs(v16)
    m = q["mat"]
    for i in range(shp[0]):
        for j in range(shp[1]):
            for ax in range(2):
                a, c, Y, Yo = (m["alpha1"], m["c11"], q["m1"] - 1j * q["nk1"] / om, q["m3"] - 1j * q["nk3"] / om) \
                    if ax == 0 else (m["alpha3"], m["c33"], q["m3"] - 1j * q["nk3"] / om, q["m1"] - 1j * q["nk1"] / om)
                Cm = np.array([[c[i, j] + a[i, j] ** 2 * m["M"][i, j], a[i, j] * m["M"][i, j]],
                               [a[i, j] * m["M"][i, j], m["M"][i, j]]], complex)
                Rm = np.array([[q["rho"][i, j], q["rf"][i, j]], [q["rf"][i, j], Y[i, j]]])
                V = np.sqrt(np.linalg.eigvals(np.linalg.solve(Rm, Cm)))
                V = V[np.argsort([-(1 / v).real ** -1 for v in V])]
                Vs = np.sqrt(m["mu_ij"][i, j] / (q["rho"][i, j] - q["rf"][i, j] ** 2 / Yo[i, j]))
                for k, Vk in enumerate((V[0], V[1], Vs)):
                    assert abs(1 / (1 / Vk).real - vel[3 * ax + k, i, j]) < 1e-10 * vel[3 * ax + k, i, j]
                    assert abs(-om * (1 / Vk).imag - att[3 * ax + k, i, j]) <= 1e-9 * abs(att[3 * ax + k, i, j])
    ref = ROOT / "oracle" / "_ref" / "libvq.so"
    if ref.exists():
        class Cx(C.Structure):
            _fields_ = [("re", C.c_double), ("im", C.c_double)]
        vq = C.CDLL(str(ref))
        for nm in ("VqP1", "VqP2", "VqS"):
            getattr(vq, nm).restype = Cx
            getattr(vq, nm).argtypes = [C.c_double] * 16
        M, a1, a3 = m["M"], m["alpha1"], m["alpha3"]
        for i in range(shp[0]):
            for j in range(shp[1]):
                args = (m["c11"][i, j] + a1[i, j] ** 2 * M[i, j], m["c13"][i, j] + a1[i, j] * a3[i, j] * M[i, j],
                        m["c33"][i, j] + a3[i, j] ** 2 * M[i, j], m["mu_ij"][i, j], a1[i, j], a3[i, j], M[i, j],
                        q["rho"][i, j], q["rf"][i, j], q["nk1"][i, j], q["nk3"][i, j], q["m1"][i, j], q["m3"][i, j], om)
                for k, (nm, lx, lz) in enumerate([("VqP1", 1., 0.), ("VqP2", 1., 0.), ("VqS", 1., 0.),
                                                   ("VqP1", 0., 1.), ("VqP2", 0., 1.), ("VqS", 0., 1.)]):
                    r = getattr(vq, nm)(lx, lz, *args)
                    # complex slowness 1/V = 1/vel - i att/omega, compared as a whole (the
                    # attenuation of the fast wave is 1e-6 of its real part)
                    s_ref = 1 / complex(r.re, r.im)
                    s = complex(1 / vel[k, i, j], -att[k, i, j] / om)
                    tol = 2e-2 if nm == "VqP2" else 1e-5
                    assert abs(s_ref - s) < tol * abs(s), (nm, lx, i, j, abs(s_ref - s) / abs(s))
    # isotropic limit (uniform eta/kappa and mu, so the staggered averages are the node values)
    iso = _pve_values(rng, 1, 8).reshape(12, 8)
    col = lambda a: np.repeat(a[None, :], 3, axis=0)
    iso[4] = 6.0
    iso[8] = 1.0
    iso[9] = 500.0
    iso[10] = 2e4            # the VTI check has no relaxation of M; the isotropic one none above Q 1e4
    zero = np.zeros((3, 8))
    v16 = [col(iso[q]) for q in range(7)] + [col(iso[7]), col(iso[7]), col(iso[8]), col(iso[9]), col(iso[9]),
                                             col(iso[10]), col(iso[11]), zero, zero]
    vel, _ = _dispersion_vti(v16, f)
    vi, _ = _dispersion([a.reshape(-1) for a in (col(iso[q]) for q in range(12))], f)
    vi = vi.reshape(3, 3, 8)
    for k, ki in ((0, 1), (1, 0), (2, 2), (3, 1), (4, 0), (5, 2)):
        assert np.allclose(vel[k], vi[ki], rtol=1e-10), k
