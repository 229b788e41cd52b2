"""The nlchains command line: argument rules, file formats and batch cadence.

CPU tests cover parsing and validation (common/main.cpp:53-134) and the
error exits (common/main.cpp:137-176); GPU tests run whole simulations and
check every file the reference writes (README.md:74-82) against the oracle.
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from parity import chain_rel_err

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "nlchains_amd" / "nlchains"


def run(*args, check=None, timeout=300):
    p = subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    if check is not None:
        assert p.returncode == check, (p.returncode, p.stdout, p.stderr)
    return p


def common(tmp, n=64, c=8, dt=0.1, b=10, extra=()):
    init = tmp / "init"
    if not init.exists():
        np.random.default_rng(0).normal(size=(c, n, 2)).tofile(init)
    return ["-n", n, "-c", c, "--dt", dt, "-b", b, "-p", tmp / "out", "-i", init, *extra]


def test_cli_lists_programs():
    p = run(check=0)
    assert "DNKG, DNLS, FPUT, Toda, dDNKG" in p.stdout


@pytest.mark.parametrize("prog,opts", [("DNKG", ["-m", "--beta"]), ("FPUT", ["--alpha", "--beta", "--fp32c"]),
                                       ("Toda", ["--alpha", "--fp32c"]), ("dDNKG", ["-m", "--beta"]),
                                       ("DNLS", ["--beta", "--no_linear_callback", "--fp32c"])])
def test_cli_help_per_program(prog, opts):
    p = run(prog, check=0)
    assert f"Options for {prog}" in p.stdout
    for o in ["--chain_length [ -n ]", "--kernel_batching [ -b ]", "--dump_interval", "--entropymask", *opts]:
        assert o in p.stdout


@pytest.mark.parametrize("extra,msg", [
    (["-b", "0"], "--kernel_batching must be positive"),
    (["--dt", "0"], "--dt and --kernel_batching must be positive"),
    (["-n", "1"], "--chain_length must be >= 2"),
    (["--dump_interval", "15"], "--dump_interval must be a multiple of --kernel_batching"),
    (["--dump_interval", "0"], "--dump_interval must be > 0"),
    (["-s", "25"], "--steps must be a multiple of --kernel_batching"),
    (["-s", "20", "-o", "20"], "--time_offset must be less than --steps"),
    (["--bogus", "1"], "unrecognised option '--bogus'"),
    (["--dt", "abc"], "is invalid"),
])
def test_cli_argument_rules(tmp_path, extra, msg):
    args = common(tmp_path)
    for k, v in zip(extra[::2], extra[1::2]):  # override or append
        if k in args:
            args[args.index(k) + 1] = v
        else:
            args += [k, v]
    p = run("FPUT", *args, "--alpha", 0, "--beta", 1, check=1)
    assert "invalid argument" in p.stderr and msg.split(" must")[0] in p.stderr, p.stderr


def test_cli_missing_required(tmp_path):
    p = run("FPUT", *common(tmp_path), "--alpha", 0, check=1)
    assert "the option '--beta' is required but missing" in p.stderr


def test_cli_io_errors(tmp_path):
    args = common(tmp_path)
    args[args.index("-i") + 1] = tmp_path / "missing"
    p = run("FPUT", *args, "--alpha", 0, "--beta", 1, check=1)
    assert "I/O error, could not read initial state" in p.stderr
    short = tmp_path / "short"
    np.zeros(10).tofile(short)
    args[args.index("-i") + 1] = short
    assert "could not read initial state" in run("FPUT", *args, "--alpha", 0, "--beta", 1, check=1).stderr
    mask = tmp_path / "mask"
    np.zeros(64, np.uint8).tofile(mask)
    p = run("FPUT", *common(tmp_path), "--alpha", 0, "--beta", 1, "--entropymask", mask, check=1)
    assert "Too few modes for entropy calculation" in p.stderr
    p = run("dDNKG", *common(tmp_path), "-m", tmp_path / "nomasses", "--beta", 1, check=1)
    assert "could not read m file" in p.stderr


def test_cli_without_gpu_fails_cleanly(tmp_path):
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        cnt = ctypes.c_int(0)
        if hip.hipGetDeviceCount(ctypes.byref(cnt)) == 0 and cnt.value > 0:
            pytest.skip("a GPU is present")
    except OSError:
        pass
    p = run("FPUT", *common(tmp_path), "--alpha", 0, "--beta", 1, check=1)
    assert "no HIP device available" in p.stderr


# ------------------------------------------------------------------ GPU

def read_entropy(prefix):
    return np.fromfile(f"{prefix}-entropy").reshape(-1, 3)


@pytest.mark.gpu
def test_cli_fput_run_files_and_cadence(gpu, tmp_path):
    n, c, dt, b = 64, 32, 0.1, 100
    # every mode carries energy (zero-energy modes make log e, hence the entropies, meaningless)
    s0 = gpu.ic_plane_waves(n, c, 0.0, 1, 5, 0.1, 1) + 0.01 * np.random.default_rng(4).normal(size=(c, n, 2))
    init = tmp_path / "init"
    s0.tofile(init)
    pre = tmp_path / "run"
    run("FPUT", "-n", n, "-c", c, "--dt", dt, "-b", b, "-s", 300, "--dump_interval", 200, "-p", pre, "-i", init,
        "--alpha", 0, "--beta", 1, check=0)
    files = sorted(p.name for p in tmp_path.iterdir() if p.name.startswith("run"))
    assert files == ["run-0", "run-200", "run-300", "run-entropy", "run-linenergies-0", "run-linenergies-200",
                     "run-linenergies-300"], files
    ent = read_entropy(pre)
    np.testing.assert_allclose(ent[:, 0], [0, 10, 20, 30])
    states = {0: s0}
    for t in (100, 200, 300):
        states[t] = O.move_real(O.FPUT, states[t - 100], 100, dt, beta=1.0)
    w = O.dispersion_real(n, 0.0)
    for row, t in zip(ent, (0, 100, 200, 300)):
        e = O.calc_linenergies(O.linenergies_real(states[t], w), 0.5 / c / n, True)
        wt, inf = O.entropies(e, True)
        assert abs(row[1] - wt) <= 1e-9 * abs(wt) and abs(row[2] - inf) <= 1e-9 * abs(inf)
        if t in (0, 200, 300):
            got = np.fromfile(f"{pre}-linenergies-{t}")
            np.testing.assert_allclose(got, e, rtol=0, atol=1e-9 * np.abs(e).max())
            dump = np.fromfile(f"{pre}-{t}").reshape(c, n, 2)
            assert np.abs(dump - states[t]).max() <= 1e-10 * np.abs(states[t]).max()
    # resume from the t=200 dump: -o 200 -s 300 reproduces the t=300 dump
    pre2 = tmp_path / "resumed"
    run("FPUT", "-n", n, "-c", c, "--dt", dt, "-b", b, "-s", 300, "-o", 200, "-p", pre2, "-i", f"{pre}-200",
        "--alpha", 0, "--beta", 1, check=0)
    assert np.array_equal(np.fromfile(f"{pre2}-300"), np.fromfile(f"{pre}-300"))
    np.testing.assert_array_equal(read_entropy(pre2)[:, 0], [20, 30])


@pytest.mark.gpu
def test_cli_autotune_run_matches_default(gpu, tmp_path):
    """--autotune times the kernel forms on the initial state (restored after)
    and may pick another wave shape: same trajectory to rounding level."""
    n, c, dt, b = 256, 40, 0.05, 50
    s0 = gpu.ic_plane_waves(n, c, 1.0, 0, 32, 0.5, 3) + 0.01 * np.random.default_rng(5).normal(size=(c, n, 2))
    init = tmp_path / "init"
    s0.tofile(init)
    common = ["-n", n, "-c", c, "--dt", dt, "-b", b, "-s", 100, "-i", init, "-m", 1, "--beta", 1]
    run("DNKG", *common, "-p", tmp_path / "plain", check=0)
    p = run("DNKG", *common, "-p", tmp_path / "tuned", "--autotune", "-v", check=0)
    assert "autotune:" in p.stderr and "*" in p.stderr
    a, t = np.fromfile(tmp_path / "plain-100"), np.fromfile(tmp_path / "tuned-100")
    assert np.abs(a - t).max() <= 1e-12 * np.abs(a).max()
    np.testing.assert_allclose(read_entropy(tmp_path / "tuned"), read_entropy(tmp_path / "plain"), rtol=1e-12)


@pytest.mark.gpu
def test_cli_entropy_limit_stops_run(gpu, tmp_path):
    n, c = 64, 8
    gpu.ic_plane_waves(n, c, 0.0, 1, 5, 0.1, 1).tofile(tmp_path / "init")
    pre = tmp_path / "lim"
    run("FPUT", "-n", n, "-c", c, "--dt", 0.1, "-b", 10, "-p", pre, "-i", tmp_path / "init", "--alpha", 0, "--beta", 1,
        "-e", 1e9, "--dump_interval", 20, check=0)
    ent = read_entropy(pre)
    assert ent.shape[0] == 1  # the entropy at t=0 is already below the limit
    assert (tmp_path / "lim-0").exists() and not (tmp_path / "lim-10").exists()


@pytest.mark.gpu
def test_cli_blow_up_stops_run(gpu, tmp_path):
    """A numerical blow-up makes the entropy non-finite and the run stops
    (common/results.cpp:64-68): Toda alpha = 1 at dt = 1, far outside the
    integrator's stability region (the oracle's chains are all non-finite
    after the first 10-step batch), asked for 2000 steps."""
    n, c, b = 64, 8, 10
    # every mode carries energy, so the entropy at t = 0 is finite
    s0 = gpu.ic_plane_waves(n, c, 0.0, 1, 5, 0.5, 1) + 0.01 * np.random.default_rng(4).normal(size=(c, n, 2))
    s0.tofile(tmp_path / "init")
    assert not np.isfinite(O.move_real(O.TODA, s0, b, 1.0, alpha=1.0)).all()
    pre = tmp_path / "boom"
    run("Toda", "-n", n, "-c", c, "--dt", 1.0, "-b", b, "-s", 2000, "--dump_interval", 2000, "-p", pre,
        "-i", tmp_path / "init", "--alpha", 1, check=0)
    ent = read_entropy(pre)
    bad = ~np.isfinite(ent[:, 1:]).all(axis=1)
    assert np.isfinite(ent[0, 1:]).all() and bad.any()
    first = int(np.argmax(bad))
    # the run ends once the batches queued at the first non-finite entropy are
    # done (two ahead, as the reference's throttle), far before step 2000
    assert first <= 3 and ent.shape[0] <= first + 3, ent[:, 0]
    assert ent[-1, 0] < 2000 * 1.0
    last = int(round(ent[-1, 0] / 1.0))
    assert (tmp_path / f"boom-{last}").exists()  # the final dump at the step the run reached


@pytest.mark.gpu
def test_cli_ddnkg_eigensystem_files_and_run(gpu, tmp_path):
    n, c = 48, 6
    masses = gpu.ic_masses(n, 0.5, 1.5, 5)
    masses.tofile(tmp_path / "m")
    om, V = gpu.eigensystem(masses + 2)
    s0 = gpu.ic_eigenmodes(n, c, om, V, 6, 0.5, 5)
    s0.tofile(tmp_path / "init")
    pre = tmp_path / "d"
    run("dDNKG", "-n", n, "-c", c, "--dt", 0.05, "-b", 20, "-s", 40, "-p", pre, "-i", tmp_path / "init", "-m",
        tmp_path / "m", "--beta", 1, check=0)
    omegas = np.fromfile(f"{pre}-omegas")
    A = np.diag(masses + 2) - np.eye(n, k=1) - np.eye(n, k=-1)
    A[0, -1] = A[-1, 0] = -1
    np.testing.assert_allclose(omegas, np.sqrt(np.linalg.eigvalsh(A)), rtol=1e-12)
    Vf = np.fromfile(f"{pre}-eigenvectors").reshape(n, n).T  # column-major on disk
    np.testing.assert_allclose(A @ Vf, Vf * omegas ** 2, atol=1e-11)
    want = O.move_real(O.DDNKG, s0, 40, 0.05, beta=1.0, mp2=masses + 2)
    got = np.fromfile(f"{pre}-40").reshape(c, n, 2)
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


@pytest.mark.gpu
def test_cli_dnls_run(gpu, tmp_path):
    n, c = 256, 4
    s0 = gpu.ic_dnls(n, c, 16, 1.0, 3)
    s0.tofile(tmp_path / "init")
    pre = tmp_path / "s"
    run("DNLS", "-n", n, "-c", c, "--dt", 0.01, "-b", 50, "-s", 100, "-p", pre, "-i", tmp_path / "init", "--beta", 1,
        check=0)
    want = O.move_dnls(s0, 100, 0.01, 1.0)
    got = np.fromfile(f"{pre}-100").reshape(c, n, 2)
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    e = O.calc_linenergies(O.linenergies_dnls(want), 1.0 / (n * c), True)
    np.testing.assert_allclose(np.fromfile(f"{pre}-linenergies-100"), e, rtol=0, atol=1e-9 * e.max())


@pytest.mark.gpu
def test_cli_dnls_fp32c(gpu, tmp_path):
    """--fp32c: the FP32-compensated step (4096-site chains) through the CLI,
    against the FP64 oracle at its stated precision; other lengths refused."""
    n, c = 4096, 3
    s0 = gpu.ic_dnls(n, c, 256, 1.0, 3)
    s0.tofile(tmp_path / "init")
    pre = tmp_path / "f"
    run("DNLS", "-n", n, "-c", c, "--dt", 0.01, "-b", 50, "-s", 100, "-p", pre, "-i", tmp_path / "init", "--beta", 1,
        "--fp32c", check=0)
    want = O.move_dnls(s0, 100, 0.01, 1.0)
    got = np.fromfile(f"{pre}-100").reshape(c, n, 2)
    assert chain_rel_err(got, want) <= 3e-6
    p = run("DNLS", "-n", 1024, "-c", c, "--dt", 0.01, "-b", 50, "-s", 100, "-p", pre, "-i", tmp_path / "init",
            "--beta", 1, "--fp32c")
    assert p.returncode != 0 and "FP32C" in (p.stdout + p.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("prog,extra,model,kw", [("FPUT", ["--alpha", 0, "--beta", 1], O.FPUT, dict(beta=1.0)),
                                                ("Toda", ["--alpha", 1], O.TODA, dict(alpha=1.0))])
def test_cli_real_fp32c(gpu, tmp_path, prog, extra, model, kw):
    """--fp32c for FPUT and Toda: the FP32-compensated step through the CLI,
    against the FP64 oracle at its stated precision; other lengths refused."""
    n, c = 256, 8
    s0 = gpu.ic_plane_waves(n, c, 0.0, 1, 32, 0.25, 3)
    s0.tofile(tmp_path / "init")
    pre = tmp_path / "f"
    run(prog, "-n", n, "-c", c, "--dt", 0.05, "-b", 50, "-s", 100, "-p", pre, "-i", tmp_path / "init", *extra,
        "--fp32c", check=0)
    want = O.move_real(model, s0, 100, 0.05, **kw)
    got = np.fromfile(f"{pre}-100").reshape(c, n, 2)
    assert chain_rel_err(got, want) <= 5e-6
    p = run(prog, "-n", 100, "-c", c, "--dt", 0.05, "-b", 50, "-s", 100, "-p", pre, "-i", tmp_path / "init", *extra,
            "--fp32c")
    assert p.returncode != 0 and "FP32C" in (p.stdout + p.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("prog,n,c,extra", [
    ("DNKG", 4096, 3, ["-m", 1, "--beta", 1]),                    # workgroup-resident kernel
    ("Toda", 3000, 2, ["--alpha", 1]),                            # workgroup kernel, ragged lanes
    ("FPUT", 100, 5, ["--alpha", 0.25, "--beta", 1, "--split_kernel"]),  # reference's --split_kernel
    ("DNKG", 20000, 2, ["-m", 1, "--beta", 1]),                  # temporally blocked kernel
])
def test_cli_long_chains_and_split_kernel(gpu, tmp_path, prog, n, c, extra):
    kw = {"DNKG": dict(m=1.0, beta=1.0), "Toda": dict(alpha=1.0), "FPUT": dict(alpha=0.25, beta=1.0)}[prog]
    s0 = 0.2 * np.random.default_rng(7).normal(size=(c, n, 2))
    s0.tofile(tmp_path / "init")
    pre = tmp_path / "r"
    p = run(prog, "-n", n, "-c", c, "--dt", 0.05, "-b", 20, "-s", 40, "-p", pre, "-i", tmp_path / "init", "-v", *extra,
            check=0)
    assert ("kernel=split" in p.stderr) == ("--split_kernel" in extra), p.stderr
    assert ("kernel=tiled" in p.stderr) == (n > 16384), p.stderr
    want = O.move_real(getattr(O, prog.upper()), s0, 40, 0.05, **kw)
    got = np.fromfile(f"{pre}-40").reshape(c, n, 2)
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    m = kw.get("m", 0.0)
    e = O.calc_linenergies(O.linenergies_real(want, O.dispersion_real(n, m)), 0.5 / c / n, m == 0.0)
    np.testing.assert_allclose(np.fromfile(f"{pre}-linenergies-40"), e, rtol=0, atol=1e-9 * e.max())


@pytest.mark.gpu
def test_cli_dnls_two_level(gpu, tmp_path):
    n, c = 16384, 2
    # every mode populated: with a band-limited start some modes come out exactly zero
    # (which ones depends on the FFT's operation order), log 0 makes the entropy
    # non-finite and the run stops at t = 0, as the reference's check_entropy does
    s0 = gpu.ic_dnls(n, c, 64, 1.0, 3) + 1e-3 * np.random.default_rng(2).normal(size=(c, n, 2))
    s0.tofile(tmp_path / "init")
    pre = tmp_path / "s"
    p = run("DNLS", "-n", n, "-c", c, "--dt", 0.01, "-b", 5, "-s", 10, "-p", pre, "-i", tmp_path / "init", "--beta", 1,
            "-v", check=0)
    assert "dnls_two_level_fft" in p.stderr
    assert read_entropy(pre).shape[0] == 3  # t = 0, 5, 10
    want = O.move_dnls(s0, 10, 0.01, 1.0)
    got = np.fromfile(f"{pre}-10").reshape(c, n, 2)
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
