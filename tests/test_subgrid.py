"""Pins helpers.subgrid, the column subsample the large-grid GPU tests check
against the float64 oracle: with the full grid's level set injected, the oracle
and the host adjoint on a subgrid reproduce the full grid at those columns."""
import numpy as np

import helpers as H
from jcm_amd import synthetic


def _case(nmember=1):
    st, fo, ge, da = synthetic.make_case(24, 12, 8, seed=3, land=True, nmember=nmember, member_t_sigma=0.5)
    q = np.array(st.specific_humidity)
    q[..., ::2, :] *= np.float32(1.1)   # condensing columns: several cloud-top levels
    return st.copy(specific_humidity=q), fo, ge, da


I = np.array([0, 5, 6, 17, 23])
J = np.array([1, 2, 7, 11])


def test_oracle_subgrid_reproduces_full_grid():
    case = _case()
    for dtype in (np.float32, np.float64):
        full, _, S = H.run_oracle(*case, dtype=dtype)
        assert bin(S).count("1") >= 2
        sub, _, S_sub = H.run_oracle(*H.subgrid(case, I, J), dtype=dtype, level_set=S)
        assert S_sub == S
        for f in ("u_wind", "v_wind", "temperature", "specific_humidity"):
            a = np.asarray(getattr(full, f))[..., I, :][..., J]
            np.testing.assert_array_equal(np.asarray(getattr(sub, f)), a, err_msg=f"{dtype.__name__} {f}")


def test_oracle_subgrid_of_a_member():
    case = _case(nmember=3)
    for m in range(3):
        full, _, S = H.run_oracle(*case, member=m)
        sub, _, _ = H.run_oracle(*H.subgrid(case, I, J, member=m), level_set=S)
        for f in ("temperature", "specific_humidity"):
            np.testing.assert_array_equal(np.asarray(getattr(sub, f)),
                                          np.asarray(getattr(full, f))[..., I, :][..., J], err_msg=f"member {m} {f}")


def test_host_adjoint_subgrid_reproduces_full_grid():
    case = _case()
    params = synthetic.draw_vjp_parameters(2)
    rng = np.random.default_rng(4)
    cot = [rng.standard_normal((8, 24, 12)).astype(np.float32) for _ in range(4)]
    _, _, S = H.run_oracle(*case, params=params)
    _, g = H.run_host_vjp(case, params, cot, level_set=S)
    _, gs = H.run_host_vjp(H.subgrid(case, I, J), params, [c[..., I, :][..., J] for c in cot], level_set=S)
    for k in ("t", "q", "u", "v", "ps", "prm"):
        a = g[k][..., I, :, :][..., J, :] if k == "prm" else g[k][..., I, :][..., J]
        np.testing.assert_array_equal(gs[k], a, err_msg=k)
