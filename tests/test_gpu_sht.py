"""GPU parity of the dynamics <-> physics transforms (libspeedy_sht.so) against
the float64 oracle (oracle/sht_oracle.py), the closed forms the mathematics
fixes, and the reference's own round trip (physics_interface_test.py:9-44).

Tolerance: the GPU computes in float32 (fp32 MFMA GEMMs for the Legendre
stages, sums of up to 427 products; two short DFT passes of n1 and n2 terms for
the Fourier stages, nlon = n1 * n2) against a float64 oracle, so each field is
compared by max|gpu - oracle| / max|oracle| <= TOL (TOL = 5e-5; float32's
epsilon is 6e-8)."""
import numpy as np
import pytest

from oracle import sht_oracle as so

pytestmark = pytest.mark.gpu

TOL = 5e-5
SIGMA = {1: [0.0, 1.0],
         3: [0.0, 0.3, 0.7, 1.0],
         7: [0.0, 0.14, 0.26, 0.42, 0.6, 0.77, 0.9, 1.0],
         8: [0.0, 0.05, 0.14, 0.26, 0.42, 0.6, 0.77, 0.9, 1.0],
         16: [0.0, 0.025, 0.05, 0.095, 0.14, 0.2, 0.26, 0.34, 0.42, 0.51, 0.6, 0.685, 0.77, 0.835, 0.9, 0.95, 1.0]}


def _rel(gpu, ref):
    gpu = gpu.detach().cpu().numpy().astype(np.float64) if hasattr(gpu, "detach") else np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    assert gpu.shape == ref.shape, (gpu.shape, ref.shape)
    scale = np.abs(ref).max()
    return float(np.abs(gpu - ref).max() / (scale if scale > 0 else 1.0))


@pytest.fixture(scope="module")
def tr():
    import torch
    from jcm_amd import transforms
    assert torch.cuda.is_available()
    transforms.load()
    return transforms


def _grids(tr, trunc):
    return tr.SpectralGrid.truncation(trunc), so.Grid.truncation(trunc)


def _random_modal(g, lead, rng, scale=1.0, lmax=None):
    c = rng.standard_normal(tuple(lead) + g.modal_shape) * scale * g.mask
    if lmax is not None:
        c = c * (np.arange(g.n_l) < lmax)
    return c


def _dynamics(tr, gg, og, kx, rng):
    orog = _random_modal(og, (), rng, 300.0, lmax=20)
    ref_t = np.linspace(220.0, 288.0, kx)
    dyn_gpu = tr.PrimitiveEquations(reference_temperature=ref_t.astype(np.float32), orography=orog.astype(np.float32),
                                    coords=tr.CoordinateSystem(gg, tr.SigmaCoordinates(SIGMA[kx])))
    dyn_ref = so.Dynamics(grid=og, sigma_boundaries=np.asarray(SIGMA[kx]),
                          reference_temperature=ref_t.astype(np.float32).astype(np.float64),
                          orography=orog.astype(np.float32).astype(np.float64))
    return dyn_gpu, dyn_ref


# every compiled Fourier factor pair: 64 = 8x8, 96 = 8x12, 128 = 8x16 (T42), 256 = 16x16,
# 320 = 16x20, 360 = 18x20, 512 = 16x32, 640 = 20x32
@pytest.mark.parametrize("trunc", [21, 31, 42, 85, 106, 119, 170, 213])
def test_to_nodal_and_to_modal_match_the_oracle(tr, trunc):
    import torch
    gg, og = _grids(tr, trunc)
    rng = np.random.default_rng(trunc)
    c = _random_modal(og, (5,), rng).astype(np.float32)
    nodal = gg.to_nodal(torch.as_tensor(c, device="cuda"))
    ref_nodal = og.to_nodal(c.astype(np.float64))
    assert _rel(nodal, ref_nodal) <= TOL
    modal = gg.to_modal(nodal)
    assert _rel(modal, c) <= TOL
    # any nodal field (not band-limited) against the oracle's analysis
    x = rng.standard_normal((3,) + og.nodal_shape).astype(np.float32)
    assert _rel(gg.to_modal(torch.as_tensor(x, device="cuda")), og.to_modal(x.astype(np.float64))) <= TOL
    # entries with m > l come out exactly zero
    assert not modal.cpu().numpy()[:, ~og.mask].any()


@pytest.mark.parametrize("dims", [(22, 23, 66, 32), (30, 31, 97, 48)])
def test_generic_fourier_stages_match_the_oracle(tr, dims):
    """Grids whose nlon is not one of the compiled factor pairs (66 = 6 x 11;
    97, prime: a single 97-point pass) take the generic LDS form of the two
    Fourier passes: the same checks as above against the oracle on that grid."""
    import torch
    gg, og = tr.SpectralGrid(*dims), so.Grid(*dims)
    rng = np.random.default_rng(dims[2])
    c = _random_modal(og, (4,), rng).astype(np.float32)
    nodal = gg.to_nodal(torch.as_tensor(c, device="cuda"))
    assert _rel(nodal, og.to_nodal(c.astype(np.float64))) <= TOL
    assert _rel(gg.to_modal(nodal), c) <= TOL
    x = rng.standard_normal((3,) + og.nodal_shape).astype(np.float32)
    assert _rel(gg.to_modal(torch.as_tensor(x, device="cuda")), og.to_modal(x.astype(np.float64))) <= TOL


def test_t425_round_trip(tr):
    """T425 (1280 x 640; its Fourier stages take the generic form, 4 latitudes a
    workgroup): a band-limited field survives to_nodal then to_modal, which the
    Gaussian quadrature and the DFT make exact up to float32 rounding."""
    import torch
    gg = tr.SpectralGrid.truncation(425)   # no float64 oracle at this size: its bases would take ~1 GB
    mode_m = np.array([0] + [m for mm in range(1, gg.n_m) for m in (mm, mm)])
    mask = mode_m[:, None] <= np.arange(gg.n_l)[None, :]
    rng = np.random.default_rng(425)
    c = (rng.standard_normal((2,) + mask.shape) * mask).astype(np.float32)
    back = gg.to_modal(gg.to_nodal(torch.as_tensor(c, device="cuda")))
    assert _rel(back, c) <= TOL


def test_repeat_calls_are_bit_identical(tr):
    import torch
    gg, og = _grids(tr, 42)
    x = torch.as_tensor(np.random.default_rng(1).standard_normal((8,) + og.nodal_shape).astype(np.float32),
                        device="cuda")
    a, b = gg.to_modal(x), gg.to_modal(x)
    assert torch.equal(a, b)
    assert torch.equal(gg.to_nodal(a), gg.to_nodal(b))


def test_solid_body_rotation_closed_form(tr):
    """vorticity 2 U0 sin(lat) / a, no divergence -> u = U0 cos(lat), v = 0."""
    import torch
    gg, og = _grids(tr, 31)
    kx, u0 = 8, 20.0
    lat = np.arcsin(og.sin_lat)[None, :]
    vor_nodal = np.broadcast_to(2.0 * u0 * np.sin(lat) / og.radius, og.nodal_shape)
    vor = np.repeat(og.to_modal(vor_nodal)[None], kx, axis=0).astype(np.float32)
    zero = np.zeros_like(vor)
    dyn, _ = _dynamics(tr, gg, og, kx, np.random.default_rng(0))
    st = tr.DynamicsState(vor, zero, zero, np.zeros((1,) + og.modal_shape, np.float32),
                          {"specific_humidity": zero})
    out = tr.dynamics_state_to_physics_state(st, dyn)
    expect = np.broadcast_to(u0 * np.cos(lat), og.nodal_shape)
    for k in range(kx):
        assert _rel(out.u_wind[k], expect) <= TOL
        assert float(out.v_wind[k].abs().max()) <= TOL * u0
    # the way back: (u, v) -> vorticity, divergence
    ds = tr.physics_state_to_dynamics_state(out, dyn)
    assert _rel(ds.vorticity, vor) <= TOL
    assert float(ds.divergence.abs().max()) <= TOL * float(np.abs(vor).max())


@pytest.mark.parametrize("trunc,kx,nsp", [(31, 8, 1), (42, 7, 1), (85, 8, 1), (31, 8, 8), (31, 16, 1), (31, 1, 1),
                                          (42, 3, 3), (213, 8, 1)])
def test_dynamics_to_physics_matches_the_oracle(tr, trunc, kx, nsp):
    gg, og = _grids(tr, trunc)
    rng = np.random.default_rng(trunc * 10 + kx)
    dyn_gpu, dyn_ref = _dynamics(tr, gg, og, kx, rng)
    f32 = lambda a: a.astype(np.float32)
    vor = f32(_random_modal(og, (kx,), rng, 1e-5))
    div = f32(_random_modal(og, (kx,), rng, 1e-6))
    tv = f32(_random_modal(og, (kx,), rng, 5.0, lmax=30))
    q = f32(_random_modal(og, (kx,), rng, 1e-3, lmax=30))
    lsp = f32(_random_modal(og, (nsp,), rng, 0.02, lmax=30))
    st = tr.DynamicsState(vor, div, tv, lsp, {"specific_humidity": q})
    out = tr.dynamics_state_to_physics_state(st, dyn_gpu)
    ref = so.dynamics_state_to_physics_state(
        so.DynamicsState(vor.astype(np.float64), div.astype(np.float64), tv.astype(np.float64),
                         lsp.astype(np.float64), {"specific_humidity": q.astype(np.float64)}), dyn_ref)
    names = ("u_wind", "v_wind", "temperature", "specific_humidity", "geopotential", "normalized_surface_pressure")
    for name, r in zip(names, ref):
        err = _rel(getattr(out, name), r)
        assert err <= TOL, (name, err)


@pytest.mark.parametrize("trunc,kx,nsp", [(31, 8, 1), (85, 8, 1), (42, 7, 7), (42, 16, 1), (42, 3, 1), (31, 1, 1)])
def test_physics_to_dynamics_matches_the_oracle(tr, trunc, kx, nsp):
    from jcm_amd.physics_interface import PhysicsState
    gg, og = _grids(tr, trunc)
    rng = np.random.default_rng(trunc + kx)
    dyn_gpu, dyn_ref = _dynamics(tr, gg, og, kx, rng)
    shape = (kx,) + og.nodal_shape
    u = (rng.standard_normal(shape) * 10.0).astype(np.float32)
    v = (rng.standard_normal(shape) * 10.0).astype(np.float32)
    t = (250.0 + rng.standard_normal(shape) * 20.0).astype(np.float32)
    q = np.abs(rng.standard_normal(shape) * 5.0).astype(np.float32)
    sp = (1.0 + rng.standard_normal(((nsp,) if nsp > 1 else ()) + og.nodal_shape) * 0.05).astype(np.float32)
    out = tr.physics_state_to_dynamics_state(PhysicsState(u, v, t, q, np.zeros(shape, np.float32), sp), dyn_gpu)
    ref = so.physics_state_to_dynamics_state(u.astype(np.float64), v.astype(np.float64), t.astype(np.float64),
                                             q.astype(np.float64), sp.astype(np.float64), dyn_ref)
    for name in ("vorticity", "divergence", "temperature_variation", "log_surface_pressure"):
        err = _rel(getattr(out, name), getattr(ref, name))
        assert err <= TOL, (name, err)
    assert _rel(out.tracers["specific_humidity"], ref.tracers["specific_humidity"]) <= TOL


@pytest.mark.parametrize("trunc,kx", [(31, 8), (85, 8), (21, 7), (31, 16), (42, 3), (31, 1)])
def test_tendency_to_dynamics_matches_the_oracle(tr, trunc, kx):
    import torch
    from jcm_amd.physics_interface import PhysicsTendency
    gg, og = _grids(tr, trunc)
    rng = np.random.default_rng(trunc * 3 + kx)
    dyn_gpu, dyn_ref = _dynamics(tr, gg, og, kx, rng)
    shape = (kx,) + og.nodal_shape
    du, dv, dt, dq = ((rng.standard_normal(shape) * s).astype(np.float32) for s in (1e-4, 1e-4, 1e-5, 1e-6))
    tend = PhysicsTendency(*(torch.as_tensor(x, device="cuda") for x in (du, dv, dt, dq)))
    out = tr.physics_tendency_to_dynamics_tendency(tend, dyn_gpu)
    ref = so.physics_tendency_to_dynamics_tendency(*(x.astype(np.float64) for x in (du, dv, dt, dq)), dyn_ref)
    for name in ("vorticity", "divergence", "temperature_variation"):
        err = _rel(getattr(out, name), getattr(ref, name))
        assert err <= TOL, (name, err)
    assert _rel(out.tracers["specific_humidity"], ref.tracers["specific_humidity"]) <= TOL
    assert tuple(out.log_surface_pressure.shape) == og.modal_shape
    assert not out.log_surface_pressure.any()
    assert out.sim_time == 0.0


def test_reference_round_trip_recovers_temperature(tr):
    """physics_interface_test.py:9-44 on the GPU: T31, 8 levels, a constant state
    (T 288, u 0.5, v -0.5, q 0.5, phi 5000, sp ones((kx, ix, il))) through
    physics_state_to_dynamics_state and back; temperature allclose."""
    import torch
    from jcm_amd.physics_interface import PhysicsState, dynamics_state_to_physics_state, \
        physics_state_to_dynamics_state
    gg, og = _grids(tr, 31)
    kx = 8
    dyn_gpu, _ = _dynamics(tr, gg, og, kx, np.random.default_rng(5))
    dyn_gpu.reference_temperature = np.full(kx, 288.0, np.float32)
    shape = (kx,) + og.nodal_shape
    ones = torch.ones(shape, device="cuda")
    state = PhysicsState(0.5 * ones, -0.5 * ones, 288.0 * ones, 0.5 * ones, 5000.0 * ones, ones.clone())
    ds = physics_state_to_dynamics_state(state, dyn_gpu)
    back = dynamics_state_to_physics_state(ds, dyn_gpu)
    assert torch.allclose(state.temperature, back.temperature)
    assert torch.allclose(state.specific_humidity, back.specific_humidity)
    assert torch.allclose(state.normalized_surface_pressure, back.normalized_surface_pressure)


def test_bad_shapes_raise(tr):
    import torch
    gg, og = _grids(tr, 21)
    with pytest.raises(ValueError):
        gg.to_modal(torch.zeros((2, 10, 10), device="cuda"))
    with pytest.raises(ValueError):
        gg.to_nodal(torch.zeros((2,) + og.nodal_shape, device="cuda"))
    dyn, _ = _dynamics(tr, gg, og, 8, np.random.default_rng(0))
    z = np.zeros((7,) + og.modal_shape, np.float32)
    with pytest.raises(ValueError):
        tr.dynamics_state_to_physics_state(tr.DynamicsState(z, z, z, z[:1], {"specific_humidity": z}), dyn)


def test_graph_replays_equal_eager_calls(tr):
    """The conversions captured in a HIP graph: several replays, with eager work and
    synchronisations between them, each equal to the eager result bit for bit."""
    import torch
    from jcm_amd.physics_interface import PhysicsTendency
    gg, og = _grids(tr, 42)
    kx = 8
    rng = np.random.default_rng(9)
    dyn, _ = _dynamics(tr, gg, og, kx, rng)
    dev = torch.device("cuda", 0)
    dyn.reference_temperature = torch.as_tensor(np.asarray(dyn.reference_temperature), device=dev)
    dyn.orography = torch.as_tensor(np.asarray(dyn.orography), device=dev)
    modal = lambda lead, sc: torch.as_tensor(_random_modal(og, lead, rng, sc).astype(np.float32), device=dev)
    st = tr.DynamicsState(modal((kx,), 1e-5), modal((kx,), 1e-6), modal((kx,), 5.0), modal((1,), 0.02),
                          {"specific_humidity": modal((kx,), 1e-3)})
    shape = (kx,) + og.nodal_shape
    tend = PhysicsTendency(*(torch.randn(shape, device=dev) * s for s in (1e-4, 1e-4, 1e-5, 1e-6)))
    fn = lambda: (tr.dynamics_state_to_physics_state(st, dyn), tr.physics_tendency_to_dynamics_tendency(tend, dyn))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        ref = fn()                       # eager, on the stream the graph is captured on
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        out = fn()
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        _ = torch.randn(1 << 20, device=dev).sum()   # eager work between replays
        torch.cuda.synchronize()
        for name in ("u_wind", "v_wind", "temperature", "specific_humidity", "geopotential",
                     "normalized_surface_pressure"):
            assert torch.equal(getattr(out[0], name), getattr(ref[0], name)), name
        for name in ("vorticity", "divergence", "temperature_variation", "log_surface_pressure"):
            assert torch.equal(getattr(out[1], name), getattr(ref[1], name)), name
        assert torch.equal(out[1].tracers["specific_humidity"], ref[1].tracers["specific_humidity"])
