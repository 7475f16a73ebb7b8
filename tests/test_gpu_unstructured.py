"""GPU parity of the unstructured triangle DVM (kbg_usolve_create, kernels k_uface / k_ucell / k_uavg)
against the oracle's orc_run_us, through the C ABI.

Tolerances: FP64 storage <= 1e-10 relative (max norm, f and w after 20 steps; observed ~1e-14);
FP32 storage <= 1e-4 relative L2 in density, velocity and temperature.
"""
import copy

import numpy as np
import pytest

from conftest import kb, oracle

pytestmark = pytest.mark.gpu


def rel_max(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rel_l2(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cfg_of(space, coll="bgk", n=16, kn=0.1, order=1):
    mk = kb.configs.cavity_2d2f2v if space == "2d2f2v" else kb.configs.cavity_2d1f2v
    return mk(nu=n, nv=n, knudsen=kn, collision=coll, interpOrder=order)


def case(space, coll="bgk", n=16, nx=12, ny=10, averaged=()):
    """A lid-driven triangulated box; `averaged` cells become cellType 3 (the mean of their first two
    neighbours), which turns their boundary faces into stored-flux faces."""
    ks = kb.SolverSet(cfg_of(space, coll, n))
    ps = kb.mesh.tri_rectangle(0.0, 1.0, nx, 0.0, 1.0, ny)
    for c in averaged:
        assert np.sum(ps.cellNeighbors[c] != -1) >= 2
        ps.cellType[c] = 3
    st = kb.mesh.init_fvm_us(ks, ps)
    return ks, ps, st, kb.mesh.timestep_us(ks, ps, st)


def boundary_cells(ps, k):
    """k boundary cells with two neighbours, two of them adjacent (so one averaged cell reads
    another: the averages run in index order)."""
    cand = [c for c in range(ps.ncell) if ps.cellType[c] == 2 and np.sum(ps.cellNeighbors[c] != -1) == 2]
    first = cand[0]
    adj = next(c for c in cand[1:] if c in ps.cellNeighbors[first])
    rest = [c for c in cand if c not in (first, adj) and not set(ps.cellNeighbors[c]) & {first, adj}][: max(0, k - 2)]
    return [first, adj] + rest


def run_both(ks, ps, st, dt, nsteps, precision):
    sg = copy.deepcopy(st)
    dev = kb.mesh.DeviceSolveUS(ks, ps, sg, precision=precision)
    rg = np.zeros(4)
    dev.step(dt, nsteps, residual=rg)
    dev.download(sg, faces=True)
    assert dev.health() == 0
    dev.close()
    ro = oracle.OracleSolveUS(ks, ps, st).run(dt, nsteps)
    return sg, st, rg, ro


@pytest.mark.parametrize("space,coll", [("2d2f2v", "bgk"), ("2d2f2v", "shakhov"), ("2d1f2v", "bgk")])
def test_us_fp64_matches_the_oracle(space, coll):
    ks, ps, st, dt = case(space, coll)
    sg, so, rg, ro = run_both(ks, ps, st, dt, 20, "fp64")
    assert np.max(np.abs(so.prim[:, 1])) > 1e-3  # the lid has driven a flow
    assert rel_max(sg.f, so.f) <= 1e-10, rel_max(sg.f, so.f)
    assert rel_max(sg.w, so.w) <= 1e-10, rel_max(sg.w, so.w)
    assert rel_max(sg.prim, so.prim) <= 1e-10
    assert rel_max(rg, ro) <= 1e-8, (rg, ro)


@pytest.mark.parametrize("space,coll,n", [("2d2f2v", "bgk", 28), ("2d2f2v", "shakhov", 16), ("2d1f2v", "bgk", 28)])
def test_us_fp32_matches_the_oracle(space, coll, n):
    ks, ps, st, dt = case(space, coll, n=n)
    sg, so, _, _ = run_both(ks, ps, st, dt, 20, "fp32")
    tol = 1e-4
    assert rel_l2(sg.prim[:, 0], so.prim[:, 0]) <= tol
    assert rel_l2(1 / sg.prim[:, 3], 1 / so.prim[:, 3]) <= tol
    assert float(np.linalg.norm(sg.prim[:, 1:3] - so.prim[:, 1:3]) / np.linalg.norm(so.prim[:, 1:3])) <= tol


def test_us_averaged_cells_and_stored_faces_match_the_oracle():
    """cellType-3 cells (update_boundary.jl:520-590) in index order, one reading another, and the
    stored-flux faces of their boundary edges (left as uploaded: here a nonzero flux)."""
    ks, ps, st, dt = case("2d2f2v", n=12, nx=8, ny=8)
    avg = boundary_cells(ps, 4)
    for c in avg:
        ps.cellType[c] = 3
    kinds = kb.mesh.face_kinds(ps)
    stored = np.flatnonzero(kinds == 2)
    assert stored.size >= 4
    rng = np.random.default_rng(5)
    st.ff[stored] = 1e-3 * rng.standard_normal(st.ff[stored].shape)
    st.fw[stored] = 1e-3 * rng.standard_normal((stored.size, 4))
    sg, so, rg, ro = run_both(ks, ps, st, dt, 12, "fp64")
    assert rel_max(sg.f, so.f) <= 1e-10
    assert rel_max(sg.w, so.w) <= 1e-10
    np.testing.assert_array_equal(sg.ff[stored], so.ff[stored])
    for c in avg:
        ids = [n for n in ps.cellNeighbors[c] if n != -1][:2]
        if any(n in avg and n > c for n in ids):
            continue  # averaged from that neighbour's value before its own average
        np.testing.assert_allclose(sg.w[c], 0.5 * (sg.w[ids[0]] + sg.w[ids[1]]), rtol=1e-15, atol=1e-300)


def test_us_phase_calls_match_the_oracle():
    """kbg_evolve then kbg_update (with residual) against orc_evolve_us / orc_update_us, phase by
    phase: face fluxes after evolve!, cells after update!."""
    ks, ps, st, dt = case("2d2f2v", "shakhov", n=12, nx=6, ny=6)
    so = copy.deepcopy(st)
    orc = oracle.OracleSolveUS(ks, ps, so)
    dev = kb.mesh.DeviceSolveUS(ks, ps, st, precision="fp64")
    for _ in range(3):
        kb._lib.check(dev.lib.kbg_reconstruct(dev.handle))  # no slopes on this path: nothing to do
        dev.evolve(dt)
        orc.evolve(dt)
        sg = copy.deepcopy(st)
        dev.download(sg, faces=True)
        assert rel_max(sg.fw, so.fw) <= 1e-12
        assert rel_max(sg.ff, so.ff) <= 1e-12
        rg, ro = np.zeros(4), np.zeros(4)
        dev.update(dt, rg)
        orc.update(dt, ro)
        dev.download(sg)
        assert rel_max(sg.f, so.f) <= 1e-12
        assert rel_max(sg.w, so.w) <= 1e-12
        assert rel_max(rg, ro) <= 1e-9
    dev.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_us_fused_step_matches_evolve_then_update(precision):
    """kbg_step (k_ustep: face fluxes evaluated inside the cell kernel, f double-buffered) against the
    phase calls kbg_evolve + kbg_update (k_uface storing the fluxes, k_ucell reading them): the same
    per-velocity flux and cell code, the face moments summed in another order (k_uface forms the
    local-frame terms as products in the storage precision, k_ustep the global-frame ones in FP64) --
    agreement to rounding: FP64 storage 1e-12 relative on w and f; FP32 storage 1e-7 on w (the FP32
    products' rounding) and 1e-6 on f, over an odd and an even number of steps (both buffer
    parities), with averaged cells and stored-flux faces."""
    ks, ps, st, dt = case("2d2f2v", "shakhov", n=16, nx=8, ny=8)
    for c in boundary_cells(ps, 4):
        ps.cellType[c] = 3
    stored = np.flatnonzero(kb.mesh.face_kinds(ps) == 2)
    rng = np.random.default_rng(9)
    st.ff[stored] = 1e-3 * rng.standard_normal(st.ff[stored].shape)
    st.fw[stored] = 1e-3 * rng.standard_normal((stored.size, 4))
    fused = kb.mesh.DeviceSolveUS(ks, ps, st, precision=precision)
    phase = kb.mesh.DeviceSolveUS(ks, ps, st, precision=precision)
    ftol, wtol = (1e-12, 1e-12) if precision == "fp64" else (1e-6, 1e-7)
    done = 0
    for n in (3, 4):
        rf, rp = np.zeros(4), np.zeros(4)
        fused.step(dt, n, residual=rf)
        for k in range(n):
            phase.evolve(dt)
            phase.update(dt, rp if k == n - 1 else None)
        done += n
        a, b = copy.deepcopy(st), copy.deepcopy(st)
        fused.download(a)
        phase.download(b)
        assert rel_max(a.f, b.f) <= ftol, rel_max(a.f, b.f)
        assert rel_max(a.w, b.w) <= wtol, rel_max(a.w, b.w)
        assert rel_max(a.prim, b.prim) <= wtol
        # the residual is built from the step's change of w, which carries w's difference amplified
        np.testing.assert_allclose(rf, rp, rtol=1e-8 if precision == "fp64" else 1e-3)
    assert done == 7 and np.max(np.abs(a.prim[:, 1])) > 1e-4
    fused.close()
    phase.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_us_mass_is_conserved_on_the_device(precision):
    """The walls carry no mass (rho_w = -SF/SG) and the two cells of a face see the same face
    moments (k_ustep sums them per face in the same order in both), so the total mass stays put: to
    FP64 rounding in FP64 storage; in FP32 storage to the FP32 rounding of the wall Maxwellian's
    amplitude (the next test closes the box to pin the interior faces alone)."""
    ks, ps, st, dt = case("2d2f2v", n=16, nx=10, ny=10)
    m0 = float(np.sum(st.w[:, 0] * ps.cellArea))
    dev = kb.mesh.DeviceSolveUS(ks, ps, st, precision=precision)
    dev.step(dt, 100)
    dev.download(st)
    dev.close()
    assert float(np.sum(st.w[:, 0] * ps.cellArea)) == pytest.approx(m0, rel=1e-12 if precision == "fp64" else 1e-8)
    assert np.max(np.abs(st.prim[:, 1])) > 1e-3


def test_us_interior_faces_conserve_mass_momentum_and_energy_exactly_in_fp32():
    """A closed box (every boundary face a stored zero flux) with a density jump: only interior
    faces move mass, momentum and energy, each leaving one cell as it enters the other, so in FP32
    storage too the totals change only by the FP64 rounding of the cell updates."""
    ks, ps, st, dt = case("2d2f2v", n=16, nx=10, ny=10)
    left = ps.cellCenter[:, 0] < 0.5
    st.f[left] *= 1.2
    st.w[left] *= 1.2
    st.prim[left, 0] *= 1.2
    fkind = kb.mesh.face_kinds(ps)
    fkind[fkind == 1] = 2
    tot0 = (st.w * ps.cellArea[:, None]).sum(axis=0)
    dev = kb.mesh.DeviceSolveUS(ks, ps, st, precision="fp32", fkind=fkind)
    dev.step(dt, 50)
    dev.download(st)
    dev.close()
    tot = (st.w * ps.cellArea[:, None]).sum(axis=0)
    assert np.max(np.abs(st.prim[:, 1])) > 1e-3  # the jump drives a flow
    assert abs(tot[0] - tot0[0]) <= 1e-13 * tot0[0]
    assert abs(tot[3] - tot0[3]) <= 1e-13 * tot0[3]
    assert np.max(np.abs(tot[1:3])) <= 1e-13 * tot0[0]


def test_us_timestep_matches_the_host_formula():
    ks, ps, st, dt = case("2d2f2v", n=16)
    dev = kb.mesh.DeviceSolveUS(ks, ps, st, precision="fp32")
    assert dev.timestep() == pytest.approx(dt, rel=1e-14)
    dts = kb.batch_timestep([dev])
    assert dts[0] == pytest.approx(dt, rel=1e-14)
    dev.close()


def test_us_in_a_mixed_batch_steps_as_alone():
    """An unstructured solve batched with a 1D sweep solve and a structured 2D solve advances exactly
    as when stepped alone (bit for bit), and so do the others."""
    ks, ps, st, dt = case("2d2f2v", n=16, nx=8, ny=8)
    k1 = kb.SolverSet(kb.configs.sod_1d1f1v())
    c1, f1 = kb.init_fvm(k1)
    k2 = kb.SolverSet(kb.configs.cavity_2d2f2v(nu=16, nv=16, nx=12, ny=12))
    c2, f2 = kb.init_fvm(k2)
    dt1, dt2 = kb.timestep(k1, c1), kb.timestep(k2, c2)
    alone = []
    for mk in (lambda: kb.mesh.DeviceSolveUS(ks, ps, st, precision="fp32"),
               lambda: kb.DeviceSolve(k1, c1, precision="fp32"), lambda: kb.DeviceSolve(k2, c2, precision="fp32")):
        alone.append(mk())
    for d, t in zip(alone, (dt, dt1, dt2)):
        d.step(t, 6)
    batch = [kb.mesh.DeviceSolveUS(ks, ps, st, precision="fp32"), kb.DeviceSolve(k1, c1, precision="fp32"),
             kb.DeviceSolve(k2, c2, precision="fp32")]
    res = np.zeros((3, 5))
    kb.batch_step(batch, [dt, dt1, dt2], 6, residuals=res)
    sa, sb = copy.deepcopy(st), copy.deepcopy(st)
    alone[0].download(sa)
    batch[0].download(sb)
    np.testing.assert_array_equal(sa.f, sb.f)
    np.testing.assert_array_equal(sa.w, sb.w)
    for a, b, c in ((alone[1], batch[1], c1), (alone[2], batch[2], c2)):
        ca, cb = copy.deepcopy(c), copy.deepcopy(c)
        a.download(ca)
        b.download(cb)
        np.testing.assert_array_equal(ca.f, cb.f)
    assert np.all(res[0, :4] > 0) and np.all(res[1, :3] > 0) and np.all(res[2, :4] > 0)
    for d in alone + batch:
        d.close()


def create_raw(ks, ps, **patch):
    """kbg_usolve_create on a descriptor whose named int32 arrays are replaced (a mesh the host
    helpers would not build)."""
    import ctypes as C

    u, keep = kb.mesh.make_udesc(ks, ps)
    for name, arr in patch.items():
        a = np.ascontiguousarray(arr, dtype=np.int32)
        keep.append(a)
        setattr(u, name, kb._lib.iptr(a))
    ctx = kb.solver.default_context()
    h = C.c_void_p()
    rc = ctx.lib.kbg_usolve_create(ctx.handle, C.byref(u), C.byref(h))
    if rc == 0:
        ctx.lib.kbg_solve_destroy(h)
    return rc, ctx.lib.kbg_last_error().decode()


def test_us_refusals():
    ks, ps, st, _ = case("2d1f2v", "shakhov", n=8, nx=3, ny=3)
    with pytest.raises(kb._lib.KbgError, match="Shakhov"):
        kb.mesh.DeviceSolveUS(ks, ps, st)
    # interpOrder 2 (the second-order unstructured reconstruct!) is not built
    ks2 = kb.SolverSet(cfg_of("2d2f2v", n=8, order=2))
    with pytest.raises(kb._lib.KbgError, match="interpOrder"):
        kb.mesh.DeviceSolveUS(ks2, ps, kb.mesh.init_fvm_us(ks2, ps))
    ks, ps, st, _ = case("2d2f2v", n=8, nx=3, ny=3)
    assert create_raw(ks, ps)[0] == 0
    cf = ps.cellFaces.copy()
    cf[2, 1] = ps.nface  # one past the last face
    rc, msg = create_raw(ks, ps, cell_faces=cf)
    assert rc < 0 and "out of range" in msg
    fc = ps.faceCells.copy()
    fc[0] = [-1, -1]  # a face with no cell
    rc, msg = create_raw(ks, ps, face_cells=fc)
    assert rc < 0 and "face 0" in msg
    lone = [c for c in range(ps.ncell) if np.sum(ps.cellNeighbors[c] != -1) == 1]
    assert lone
    ct = ps.cellType.copy()
    ct[lone[0]] = 3
    rc, msg = create_raw(ks, ps, cell_type=ct)
    assert rc < 0 and "two neighbours" in msg
    dev = kb.mesh.DeviceSolveUS(ks, ps, st)
    with pytest.raises(kb._lib.KbgError, match="slopes"):
        kb._lib.check(dev.lib.kbg_upload_slopes(dev.handle, None, None))
    dev.close()
