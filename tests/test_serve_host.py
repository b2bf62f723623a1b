"""Host-side request checks of the serving loop (no GPU): invalid requests are recognised
before they could reach a packed job, and the summary counts errors and unconverged
solves apart from the served ones."""
import numpy as np

from dg_nn_public_amd import meshgen
from dg_nn_public_amd.serve import Request, Result, summarize, validate


def _req(D=3, n=2, **over):
    m = meshgen.cube_mesh(n) if D == 3 else meshgen.square_mesh(n)
    p = meshgen.sine_problem(m, sigma=1.0)
    kw = dict(rid=0, mesh=m, kappa=p.kappa, f=p.f, u_bc=p.u_bc, arrival=0.0)
    kw.update(over)
    return Request(**kw)


def test_valid_request_passes():
    assert validate(_req()) is None and validate(_req(D=2, n=3)) is None


def test_zero_nonfinite_or_short_diffusivity_is_refused():
    r = _req()
    k = r.kappa.copy()
    k[3] = 0.0
    assert "diffusivity" in validate(_req(kappa=k))
    k[3] = np.nan
    assert "diffusivity" in validate(_req(kappa=k))
    assert "diffusivity" in validate(_req(kappa=r.kappa[:-1]))


def test_nonfinite_source_or_boundary_data_is_refused():
    r = _req()
    f = r.f.copy()
    f[7] = np.inf
    assert "source" in validate(_req(f=f))
    u = r.u_bc.copy()
    u[0] = np.nan
    assert "boundary" in validate(_req(u_bc=u))
    assert "nele*nloc" in validate(_req(f=r.f[:-2]))


def test_malformed_mesh_is_refused():
    r = _req()
    bad = meshgen.Mesh(D=3, x_ref_in=r.mesh.x_ref_in, cells=r.mesh.cells + 10 ** 6,
                       verts=r.mesh.verts)
    assert "out of range" in validate(_req(mesh=bad))
    bad2 = meshgen.Mesh(D=3, x_ref_in=r.mesh.x_ref_in[:, :2], cells=r.mesh.cells,
                        verts=r.mesh.verts)
    assert "malformed" in validate(_req(mesh=bad2))


def test_summary_separates_errors_and_unconverged():
    res = [Result(0, 0.0, 0.0, 1.0, 10, 1e-11, 0),
           Result(1, 0.5, 0.5, 2.0, 400, 3e-9, 0, converged=False),
           Result(2, 0.5, 0.5, 0.5, 0, float("nan"), -1, converged=False, error="bad")]
    s = summarize(res)
    assert s["requests"] == 3 and s["errors"] == 1 and s["not_converged"] == 1
    assert s["jobs"] == 1 and abs(s["solves_per_s"] - 2 / 2.0) < 1e-12
