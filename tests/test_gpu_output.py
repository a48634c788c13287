"""Output files of the FTLE path (the reference's write_lcs and
write_flowmap_substep), written from the GPU results.

Checked against an independent reading of the reference's I/O rules:
  * file names ./cfd2lcs_output/<label>_<index>.dat with index nint(t/h),
    minus nint(T/h) for FTLE_FWD, and ten zero-padded digits for 1000..9999
    (io_nohdf5_m.f90:64-118);
  * the Tecplot header of write_ftle (io_nohdf5_m.f90:205-238) and its
    variable list under WRITE_FLOWMAP / WRITE_VELOCITY / WRITE_BCFLAG;
  * BLOCK data in ES14.6, ten per line (write_data_from_master,
    io_nohdf5_m.f90:267-332), every field the Fortran edit of the library's
    own in-memory result (FTLE, composed flowmap) or of the inputs (grid,
    flags, velocity);
  * the velocity columns: lp%ugrid after reconstruct_flowmap, i.e. u at the
    output time for FTLE_BKWD and the velocity stored at slot step+1 for
    FTLE_FWD, zero when that slot was never written (flowmap_m.f90:212-247);
  * the slot files ./cfd2lcs_tmp/0000_<label>-LP-FM_<step>.dat and
    0000_U_NP1_<step>.dat as unformatted records (flowmap_m.f90:106-156),
    against the closed-form h-interval displacement of a spatially uniform,
    linearly time-varying velocity (RK2 is exact for it).
"""
import math
import os

import numpy as np
import pytest

import flows as F
from fortran_io import es, read_tecplot, read_unformatted

pytestmark = pytest.mark.gpu


def _nint(x):
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def _grid():
    L = 2.0 * np.pi
    X, Y, Z = F.cell_grid((11, 7, 5), (0, 0, 0), (L, L, L))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    fl[0, :, :] = F.LCS_WALL
    fl[-1, :, :] = F.LCS_WALL
    return X, Y, Z, fl


def _vel(X, Y, Z, t):
    u, v, w = F.abc_velocity(X, Y, Z)
    s = 1.0 + 0.5 * math.sin(3.0 * t)
    return u * s, v * s, w * (2.0 - s)


def _fields(a, dtype):
    return [es(x) for x in np.asarray(a, dtype=dtype).ravel(order="F")]


def test_tecplot_files_dp(tmp_path, monkeypatch):
    import libcfd2lcs_amd as L
    monkeypatch.chdir(tmp_path)
    lib = L.Cfd2lcs("dp", write_output=True)
    lib.finalize()
    X, Y, Z, fl = _grid()
    lib.init(X, Y, Z, fl)
    T, h, dt = 0.6, 0.2, 0.05
    hf = lib.diagnostic_init(L.FTLE_FWD, 0, T, h, "fw")
    hb = lib.diagnostic_init(L.FTLE_BKWD, 0, T, h, "bk")
    lib.set_option("INTEGRATOR", L.RK2)
    lib.set_option("INTERPOLATOR", L.LINEAR)
    lib.set_option("WRITE_BCFLAG", L.LCS_TRUE)
    ring = {}
    expect = []
    for s in range(30):
        t = s * dt
        lib.update(*_vel(X, Y, Z, t), t)
        assert lib.error == 0
        for hd, label, fwd in ((hf, "fw", True), (hb, "bk", False)):
            if lib.noutput(hd) == len([e for e in expect if e["label"] == label]):
                continue
            step = _nint(math.fmod(t, T) / h)
            ring[step] = t
            if fwd:
                tu = ring.get(step + 1)
                uvw = _vel(X, Y, Z, tu) if tu is not None else [np.zeros_like(X)] * 3
            else:
                uvw = _vel(X, Y, Z, t)
            idx = _nint(t / h) - (_nint(T / h) if fwd else 0)
            expect.append({"label": label, "t": t, "idx": idx, "ftle": lib.ftle(hd), "fm": lib.composed(hd),
                           "u": uvw})
    assert lib.flush_output() == 0
    assert len(expect) == 14
    assert any(e["label"] == "fw" and not np.any(e["u"][0]) for e in expect), "the zero-velocity case must occur"
    names = ["X", "Y", "Z", "FM-X", "FM-Y", "FM-Z", "U-X", "U-Y", "U-Z", "FLAG", "FTLE"]
    for e in expect:
        path = os.path.join("cfd2lcs_output", f"{e['label']}_{e['idx']}.dat")
        assert os.path.exists(path), path
        header, vnames, data, dims = read_tecplot(path)
        assert vnames == names
        assert dims == [11, 7, 5]
        lab = e["label"] + "-FTLE"
        assert header == ([f'TITLE = "{lab}"', "VARIABLES ="] + [f'"{n}"' for n in names] +
                          ["ZONE", f'T = "{lab}"', "STRANDID = 0000000001, SOLUTIONTIME = " + es(e["t"], 18, 4),
                           "I = 0000000011 J = 0000000007 K = 0000000005",
                           "ZONETYPE = Ordered, DATAPACKING = BLOCK"])
        want = {"X": X, "Y": Y, "Z": Z, "FM-X": e["fm"][0], "FM-Y": e["fm"][1], "FM-Z": e["fm"][2],
                "U-X": e["u"][0], "U-Y": e["u"][1], "U-Z": e["u"][2], "FLAG": fl.astype(np.float64),
                "FTLE": e["ftle"]}
        for n in names:
            assert data[n] == _fields(want[n], np.float64), f"{path}: column {n}"
    assert not [f for f in os.listdir("cfd2lcs_output") if f.endswith(".part")]
    lib.finalize()


def test_tecplot_variable_switches_and_slots(tmp_path, monkeypatch):
    import libcfd2lcs_amd as L
    monkeypatch.chdir(tmp_path)
    lib = L.Cfd2lcs("dp", write_output=True)
    lib.finalize()
    X, Y, Z, fl = _grid()
    fl[:] = F.LCS_INTERNAL
    lib.init(X, Y, Z, fl)
    T, h, dt = 0.4, 0.2, 0.05
    hf = lib.diagnostic_init(L.FTLE_FWD, 0, T, h, "fw")
    hb = lib.diagnostic_init(L.FTLE_BKWD, 0, T, h, "bk")
    lib.set_option("INTEGRATOR", L.RK2)
    lib.set_option("INTERPOLATOR", L.LINEAR)
    lib.set_option("WRITE_FLOWMAP", L.LCS_FALSE)
    lib.set_option("WRITE_SLOTS", L.LCS_TRUE)
    c0 = np.array([0.3, -0.2, 0.1])
    outs = []
    for s in range(10):
        t = s * dt
        u, v, w = (np.full(X.shape, c0[d] * (1.0 + t)) for d in range(3))
        lib.update(u, v, w, t)
        assert lib.error == 0
        if lib.noutput(hb) > len(outs):
            outs.append(t)
    assert lib.flush_output() == 0
    assert len(outs) == 2
    npt = X.size
    for t in outs:
        step = _nint(math.fmod(t, T) / h)
        # exact h-interval displacement of c0 (1 + s) over [t - h, t]
        disp = c0 * (h + (t * t - (t - h) * (t - h)) / 2.0)
        for label, sign in (("fw", 1.0), ("bk", -1.0)):
            recs = read_unformatted(os.path.join("cfd2lcs_tmp", f"0000_{label}-LP-FM_{step:04d}.dat"), np.float64)
            assert [r.size for r in recs] == [npt] * 3
            for d in range(3):
                assert np.abs(recs[d] - sign * disp[d]).max() <= 1e-13, (label, t, d)
        recs = read_unformatted(os.path.join("cfd2lcs_tmp", f"0000_U_NP1_{step:04d}.dat"), np.float64)
        for d in range(3):
            assert np.array_equal(recs[d], np.full(npt, c0[d] * (1.0 + t)))
        header, vnames, data, _ = read_tecplot(os.path.join("cfd2lcs_output", f"bk_{_nint(t / h)}.dat"))
        assert vnames == ["X", "Y", "Z", "U-X", "U-Y", "U-Z", "FTLE"]
    lib.finalize()
    # without velocity columns or flowmap columns
    lib.init(X, Y, Z, fl)
    hb = lib.diagnostic_init(L.FTLE_BKWD, 0, T, h, "plain")
    lib.set_option("WRITE_FLOWMAP", L.LCS_FALSE)
    lib.set_option("WRITE_VELOCITY", L.LCS_FALSE)
    for s in range(6):
        lib.update(*_vel(X, Y, Z, s * dt), s * dt)
    assert lib.flush_output() == 0 and lib.error == 0
    _, vnames, data, _ = read_tecplot(os.path.join("cfd2lcs_output", "plain_1.dat"))
    assert vnames == ["X", "Y", "Z", "FTLE"]
    lib.finalize()


def test_tecplot_sp_and_index_padding(tmp_path, monkeypatch):
    """Single-precision library (values written as REAL(4)), and the
    reference's i10.10 file index for 1000..9999."""
    import libcfd2lcs_amd as L
    monkeypatch.chdir(tmp_path)
    lib = L.Cfd2lcs("sp", write_output=True)
    lib.finalize()
    X, Y, Z, fl = F.abc_grid((6, 5, 4))
    lib.init(X, Y, Z, fl)
    hb = lib.diagnostic_init(L.FTLE_BKWD, 0, 0.002, 0.001, "q")
    got = []
    times = (1.0, 1.001, 1.002, 1.003, 1.004)
    h32 = np.float32(0.001)
    for s, t in enumerate(times):
        lib.update(*_vel(X, Y, Z, t), t)
        assert lib.error == 0
        t32 = np.float32(t)
        # the trigger int(t_np1/h) /= int(t_n/h) in REAL(4) (cfd2lcs.f90:147):
        # 1.002/0.001 rounds to 1001.99994, so that update has no output
        fires = s > 0 and int(t32 / h32) != int(np.float32(times[s - 1]) / h32)
        assert lib.noutput(hb) == len(got) + int(fires), t
        if fires:
            got.append((_nint(float(np.float32(t32 / h32))), t32, lib.ftle(hb)))
    assert lib.flush_output() == 0
    assert len(got) == 2
    for idx, t32, ftle in got:
        assert 1000 <= idx <= 9999
        path = os.path.join("cfd2lcs_output", f"q_{idx:010d}.dat")
        header, vnames, data, _ = read_tecplot(path)
        assert header[-3] == "STRANDID = 0000000001, SOLUTIONTIME = " + es(float(t32), 18, 4)
        assert data["FTLE"] == _fields(ftle, np.float32)
        assert data["X"] == _fields(X.astype(np.float32), np.float32)
    lib.finalize()


def test_velocity_invariant_columns(tmp_path, monkeypatch):
    """VELOCITY_INVARIANTS (invariants_m.f90:15-198): Q, C, D, H, L2 columns
    after FLAG, computed on the GPU from the same velocity at T0 as the U
    columns, against the oracle's invariants of that velocity (zero field
    when the forward diagnostic's slot step+1 was never written)."""
    import libcfd2lcs_amd as L
    from oracle_api import Oracle
    monkeypatch.chdir(tmp_path)
    lib = L.Cfd2lcs("dp", write_output=True)
    lib.finalize()
    X, Y, Z, fl = _grid()
    o = Oracle(X, Y, Z, fl)
    lib.init(X, Y, Z, fl)
    T, h, dt = 0.6, 0.2, 0.05
    hf = lib.diagnostic_init(L.FTLE_FWD, 0, T, h, "fw")
    hb = lib.diagnostic_init(L.FTLE_BKWD, 0, T, h, "bk")
    lib.set_option("VELOCITY_INVARIANTS", L.LCS_TRUE)
    lib.set_option("WRITE_FLOWMAP", L.LCS_FALSE)
    ring, expect = {}, []
    for s in range(18):
        t = s * dt
        lib.update(*_vel(X, Y, Z, t), t)
        assert lib.error == 0
        for hd, label, fwd in ((hf, "fw", True), (hb, "bk", False)):
            if lib.noutput(hd) == len([e for e in expect if e[0] == label]):
                continue
            step = _nint(math.fmod(t, T) / h)
            ring[step] = t
            tu = ring.get(step + 1) if fwd else t
            inv = o.invariants(*_vel(X, Y, Z, tu)) if tu is not None else o.invariants()
            idx = _nint(t / h) - (_nint(T / h) if fwd else 0)
            expect.append((label, idx, inv))
    assert lib.flush_output() == 0
    assert len(expect) == 8
    for label, idx, inv in expect:
        _, vnames, data, _ = read_tecplot(os.path.join("cfd2lcs_output", f"{label}_{idx}.dat"))
        assert vnames == ["X", "Y", "Z", "U-X", "U-Y", "U-Z", "Q", "C", "D", "H", "L2", "FTLE"]
        for name, ref in zip(("Q", "C", "D", "L2"), (inv[0], inv[1], inv[2], inv[4])):
            assert data[name] == _fields(ref, np.float64), (label, idx, name)
        got_h = np.array([float(x) for x in data["H"]])
        ref_h = inv[3].ravel(order="F")
        assert np.array_equal(np.isnan(got_h), np.isnan(ref_h))
        m = np.isfinite(ref_h)
        assert np.all(np.abs(got_h[m] - ref_h[m]) <= 1e-6 * np.abs(ref_h[m])), (label, idx)
    o.close()
    lib.finalize()


def test_output_directory_error(tmp_path, monkeypatch):
    """A path that cannot be a directory sets the sticky error flag, as the
    reference's failed mkdir does (cfd2lcs.f90:35-40)."""
    import libcfd2lcs_amd as L
    monkeypatch.chdir(tmp_path)
    with open("cfd2lcs_output", "w") as f:
        f.write("not a directory\n")
    lib = L.Cfd2lcs("dp", write_output=True)
    lib.finalize()
    X, Y, Z, fl = F.abc_grid(6)
    lib.init(X, Y, Z, fl)
    lib.diagnostic_init(L.FTLE_BKWD, 0, 0.2, 0.1, "e")
    for s in range(4):
        lib.update(*_vel(X, Y, Z, 0.05 * s), 0.05 * s)
    assert lib.error == 1
    lib.finalize()
    assert lib.error == 0


def test_refined_aux_grid_files(tmp_path, monkeypatch):
    """resolution = 1 with AUX_GRID: the FTLE file is written on the refined
    grid (I/J/K and X/Y/Z of new_sgrid_from_sgrid), its velocity columns are
    the velocity at T0 at the refined nodes (exact here: the field is uniform
    in space), every set writes its slot file <label>-LP-FM / -LP0..-LP5-FM
    on the refined grid with the exact uniform displacement, and u_np1 is
    written once on the flow grid.  The grid is isotropic: AUX_GRID shifts a
    node by a fifth of its mean face-neighbour distance, which on an
    anisotropic grid differs at the boundary nodes and leaves the auxiliary
    grids non-rectilinear (rejected, as the reference's non-rectilinear path
    is not built)."""
    import libcfd2lcs_amd as L
    monkeypatch.chdir(tmp_path)
    lib = L.Cfd2lcs("dp", write_output=True)
    lib.finalize()
    L2 = 2.0 * np.pi
    X, Y, Z = F.cell_grid((8, 6, 5), (0, 0, 0), (L2, L2 * 6 / 8, L2 * 5 / 8))
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    lib.init(X, Y, Z, fl)
    lib.set_option("AUX_GRID", L.LCS_TRUE)
    T, h, dt = 0.4, 0.2, 0.05
    hb = lib.diagnostic_init(L.FTLE_BKWD, 1, T, h, "rb")
    lib.set_option("INTEGRATOR", L.RK2)
    lib.set_option("INTERPOLATOR", L.LINEAR)
    lib.set_option("WRITE_SLOTS", L.LCS_TRUE)
    shape, xyz = lib.lcs_grid(hb)
    assert shape == (16, 12, 10)
    c0 = np.array([0.3, -0.2, 0.1])
    outs = []
    for s in range(10):
        t = s * dt
        u, v, w = (np.full(X.shape, c0[d] * (1.0 + t)) for d in range(3))
        lib.update(u, v, w, t)
        assert lib.error == 0
        if lib.noutput(hb) > len(outs):
            outs.append(t)
    assert lib.flush_output() == 0
    assert len(outs) == 2
    nl = int(np.prod(shape))
    for t in outs:
        step = _nint(math.fmod(t, T) / h)
        disp = c0 * (h + (t * t - (t - h) * (t - h)) / 2.0)
        for suffix in ["-LP"] + [f"-LP{a}" for a in range(6)]:
            recs = read_unformatted(os.path.join("cfd2lcs_tmp", f"0000_rb{suffix}-FM_{step:04d}.dat"), np.float64)
            assert [r.size for r in recs] == [nl] * 3, suffix
            for d in range(3):
                assert np.abs(recs[d] + disp[d]).max() <= 1e-13, (suffix, t, d)
        recs = read_unformatted(os.path.join("cfd2lcs_tmp", f"0000_U_NP1_{step:04d}.dat"), np.float64)
        assert [r.size for r in recs] == [X.size] * 3
        header, vnames, data, dims = read_tecplot(os.path.join("cfd2lcs_output", f"rb_{_nint(t / h)}.dat"))
        assert vnames == ["X", "Y", "Z", "FM-X", "FM-Y", "FM-Z", "U-X", "U-Y", "U-Z", "FTLE"]
        assert "I = 0000000016 J = 0000000012 K = 0000000010" in header
        for d, n in enumerate("XYZ"):
            assert data[n] == _fields(xyz[d], np.float64)
            assert data["U-" + n] == _fields(np.full(nl, c0[d] * (1.0 + t)), np.float64)
        assert dims == [16, 12, 10]
        if t == outs[-1]:
            assert data["FTLE"] == _fields(lib.ftle(hb), np.float64)
    lib.finalize()
