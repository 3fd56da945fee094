"""zand run directories on the device engine: the _save/ files carry the same
times and currents as the engine's own adaptive loop, a checkpoint resumes the
trajectory, and the command line runs an Arrays/ directory end to end."""
import os
import sys

import numpy as np
import pytest

from zandpack_amd import builder, rundir

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem():
    p = builder.config_small(NK=2, seed=3)
    p.pulse = builder.Pulse("step", amp=0.2, p1=0.0, weights=(0.5, -0.5))
    return p


T1, H0, EPS = 0.6, 0.01, 1e-8


def _reference_run(p):
    from zandpack_amd.engine import Propagator
    prop = Propagator(p, "fp64")
    r = prop.propagate(0.0, T1, H0, EPS)
    prop.close()
    return r


def test_save_dir_matches_engine_loop(native, tmp_path):
    from zandpack_amd.engine import Propagator
    p = _problem()
    say("rundir save", p.dims)
    ref = _reference_run(p)
    prop = Propagator(p, "fp64")
    d = str(tmp_path / "x_save")
    w = rundir.SaveWriter(d, p.dims[2], saveevery=5, n_dm_compress=1, t0=0.0, sig0=prop.sigma())
    t, h, acc, rej = rundir.propagate_to_save(prop, w, 0.0, T1, H0, EPS)
    sig_end = prop.sigma()
    prop.close()
    assert acc == ref["n_steps"] and rej == ref["n_rejected"]
    assert abs(h - ref["h_final"]) <= 1e-10 * abs(ref["h_final"])
    T, C = rundir.combine_currents([d], p.dims[2])
    k = len(T)
    assert k == (acc // 5) * 5 and k > 0
    np.testing.assert_allclose(T, ref["times"][:k], rtol=1e-10, atol=0)
    for a in range(p.dims[2]):
        np.testing.assert_allclose(C[a], ref["J"][:k, a], rtol=1e-10, atol=1e-14)
    last = max(int(f[2:-4]) for f in os.listdir(d) if f.startswith("DM") and not f.startswith("DMt"))
    dm = np.load(os.path.join(d, f"DM{last}.npy"))
    assert dm.dtype == np.complex64 and dm.shape[1:] == sig_end.shape
    if last == acc:
        np.testing.assert_allclose(dm[-1], sig_end, atol=1e-6)


def test_checkpoint_resume_repeats_trajectory(native, tmp_path):
    from zandpack_amd.engine import Propagator
    p = _problem()
    say("rundir checkpoint", p.dims)
    ref = _reference_run(p)
    d = str(tmp_path / "x_save")
    prop = Propagator(p, "fp64")
    w = rundir.SaveWriter(d, p.dims[2], saveevery=1000, t0=0.0, sig0=prop.sigma())
    rundir.propagate_to_save(prop, w, 0.0, T1, H0, EPS, checkpoints=(0.3,))
    prop.close()
    t_ck, h_ck, s, ps, om = rundir.load_checkpoint(d)
    i = int(np.argmin(np.abs(ref["times"] - t_ck)))
    assert abs(ref["times"][i] - t_ck) <= 1e-12 and t_ck > 0.3 and ref["times"][i - 1] <= 0.3
    prop = Propagator(p, "fp64", state=(s, ps, om))
    r = prop.propagate(t_ck, T1, h_ck, EPS)
    prop.close()
    n = len(ref["times"]) - i
    assert r["n_steps"] == n - 1
    np.testing.assert_allclose(r["times"], ref["times"][i:], rtol=1e-10, atol=0)
    np.testing.assert_allclose(r["J"], ref["J"][i:], rtol=1e-9, atol=1e-13)


def test_command_line_on_arrays_dir(native, tmp_path):
    p = _problem()
    say("rundir cli", p.dims)
    ref = _reference_run(p)
    rundir.write_arrays(p, str(tmp_path / "x" / "Arrays"))
    rc = rundir.main([f"Dir={tmp_path}", "name=x", f"t1={T1}", f"stepsize={H0}", f"eps={EPS}",
                      "saveevery=4", "pulse=step", "amp=0.2", "p1=0", "weights=0.5,-0.5"])
    assert rc == 0
    T, C = rundir.combine_currents([str(tmp_path / "x_save")], p.dims[2])
    k = len(T)
    assert k > 0
    np.testing.assert_allclose(T, ref["times"][:k], rtol=1e-10, atol=0)
    np.testing.assert_allclose(C[0], ref["J"][:k, 0], rtol=1e-9, atol=1e-13)


def test_command_line_reads_initial_py(native, tmp_path):
    """The run's settings come from its Initial.py (read, not imported); only the bias
    model is given as keywords."""
    p = _problem()
    say("rundir cli Initial.py", p.dims)
    ref = _reference_run(p)
    rundir.write_arrays(p, str(tmp_path / "x" / "Arrays"))
    (tmp_path / "Initial.py").write_text(
        "import numpy as np\n"
        f"name = 'x'\neps = {EPS!r}\nt0, t1 = 0.0, {T1!r}\nstepsize = {H0!r}\n"
        "saveevery = 4\nusesave = False\ncheckpoints = [0.3]\nsave_checkpoints = False\n")
    rc = rundir.main([f"Dir={tmp_path}", "pulse=step", "amp=0.2", "p1=0", "weights=0.5,-0.5"])
    assert rc == 0
    T, C = rundir.combine_currents([str(tmp_path / "x_save")], p.dims[2])
    k = len(T)
    assert k > 0
    np.testing.assert_allclose(T, ref["times"][:k], rtol=1e-10, atol=0)
    np.testing.assert_allclose(C[0], ref["J"][:k, 0], rtol=1e-9, atol=1e-13)
    assert not any(f.startswith("last_") for f in os.listdir(tmp_path / "x_save"))


def test_command_line_nonorthogonal(native, tmp_path):
    """A directory with S^(-0.5) run as nozand runs it."""
    from zandpack_amd.engine import Propagator
    p = _problem()
    NK, n = p.dims[:2]
    Sm05 = builder.lowdin_factors(builder.random_overlap(NK, n, seed=8))
    q = builder.to_nonorthogonal(p, Sm05)
    say("rundir cli nonorthogonal", q.dims)
    prop = Propagator(q, "fp64")
    ref = prop.propagate(0.0, T1, H0, EPS)
    prop.close()
    rundir.write_arrays(p, str(tmp_path / "x" / "Arrays"), Sm05=Sm05)
    rc = rundir.main([f"Dir={tmp_path}", "name=x", f"t1={T1}", f"stepsize={H0}", f"eps={EPS}",
                      "saveevery=4", "pulse=step", "amp=0.2", "p1=0", "weights=0.5,-0.5",
                      "nonorthogonal=True"])
    assert rc == 0
    T, C = rundir.combine_currents([str(tmp_path / "x_save")], p.dims[2])
    k = len(T)
    assert k > 0
    np.testing.assert_allclose(T, ref["times"][:k], rtol=1e-10, atol=0)
    np.testing.assert_allclose(C[1], ref["J"][:k, 1], rtol=1e-9, atol=1e-13)


@pytest.mark.parametrize("nonortho", [False, True])
def test_pi_output(native, tmp_path, nonortho):
    """Per-lead Π of the resident state (zand's Q_np / nozand's PI_np) against the
    oracle, and the PIa files of a save_PI run."""
    from oracle import zand_oracle as O
    from zandpack_amd.engine import Propagator
    p = _problem()
    if nonortho:
        NK, n = p.dims[:2]
        p = builder.to_nonorthogonal(p, builder.lowdin_factors(builder.random_overlap(NK, n, seed=2)))
    rng = np.random.default_rng(4)
    s0, ps0, om0 = p.zero_state()
    ps0 = 0.05 * (rng.normal(size=ps0.shape) + 1j * rng.normal(size=ps0.shape))
    say("pi output", nonortho)
    prop = Propagator(p, "fp64", state=(s0, ps0, om0))
    NK, n, NL, NX, NOT = p.dims
    P = ps0.reshape(NK, NL, NX * NOT, n)
    X = p.Ixi.reshape(NK, NL, NX * NOT, n)
    r = np.einsum("kamo,kamp->kaop", P, X) / p.hbar
    want = r if nonortho else r + r.conj().transpose(0, 1, 3, 2)
    got = prop.pi(hermitian=not nonortho)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    if not nonortho:
        _, Q = O.rhs(p.consts(), s0, ps0, om0, p.H0, np.zeros(NL), return_pi=True)
        assert np.abs(got - Q).max() < 1e-12 * np.abs(Q).max()
    d = str(tmp_path / "x_save")
    w = rundir.SaveWriter(d, NL, saveevery=3, t0=0.0, sig0=prop.sigma(), save_pi=True)
    rundir.propagate_to_save(prop, w, 0.0, 0.05, 0.01, 1e-8)
    pia = np.load(os.path.join(d, "PIa3.npy"))
    assert pia.shape == (3, NK, NL, n, n) and pia.dtype == np.complex64
    assert np.abs(pia[0] - want).max() < 1e-6 * np.abs(want).max()
    prop.close()
