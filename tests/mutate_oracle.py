"""Mutation check of the CPU oracle (run by hand, on the CPU; pytest does not
collect this file): each entry plants one plausible mistake in a temporary
copy of oracle/fastgb.py and runs the oracle's tests there; every mutation
must make at least one of them fail.

    python tests/mutate_oracle.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MUTS = {
    "light-travel phase sign (om*kdotP -> -om*kdotP)": ("    kdotP = om[:, None, None] * kdotP - argS", "    kdotP = -om[:, None, None] * kdotP - argS"),
    "TDI delay factor sign (exp(-i omL) -> exp(+i omL))": ("    fctr = np.exp(-1.0j * omL)", "    fctr = np.exp(1.0j * omL)"),
    "get_phi beta (drop /e)": ("        beta = (1.0 - np.sqrt(1.0 - ea * ea)) / ea", "        beta = (1.0 - np.sqrt(1.0 - ea * ea))"),
    "get_phi beta (sign of sqrt)": ("        beta = (1.0 - np.sqrt(1.0 - ea * ea)) / ea", "        beta = (1.0 + np.sqrt(1.0 - ea * ea)) / ea"),
    "link argument (1 + k.r -> 1 - k.r)": ("        arg_ij = 0.5 * fonfs[:, s, :] * (1 + kdotr[ij])", "        arg_ij = 0.5 * fonfs[:, s, :] * (1 - kdotr[ij])"),
    "transfer phase (arg + kP -> arg - kP)": ("np.exp(-1.0j * (arg_ij + kdotP[:, s]))", "np.exp(-1.0j * (arg_ij - kdotP[:, s]))"),
    "X combination index (G21 -> G12)": ('    Xsl = Gs["21"] - Gs["31"]', '    Xsl = Gs["12"] - Gs["31"]'),
    "fdot phase term dropped": ("        + np.pi * fdot[:, None, None] * (xi**2)\n", "\n"),
    "fftshift dropped": ("    XYZf = np.fft.fftshift(XYZf_slow, axes=-1)", "    XYZf = XYZf_slow"),
    "transform scale 0.5T/N -> T/N": ("    fctr = 0.5 * T / N\n", "    fctr = T / N\n"),
    "start bin offset (q - M/2 -> q - M/2 + 1)": ("    fmin = (q - M / 2) / T", "    fmin = (q - M / 2 + 1) / T"),
    "fft sign (fft -> ifft*M)": ("np.fft.fft(XYZsl, axis=-1)", "(XYZsl.shape[-1] * np.fft.ifft(XYZsl, axis=-1))"),
    "cross antenna factor sign (DC uses -Aplus -> +Aplus)": ("    DC = -Aplus * sinps - 1.0j * Across * cosps", "    DC = Aplus * sinps - 1.0j * Across * cosps"),
}
TESTS = ["tests/test_oracle_physics.py", "tests/test_oracle.py"]
res = []
for name, (a, b) in MUTS.items():
    d = tempfile.mkdtemp(prefix="mut_", dir="/tmp")
    for sub in ("oracle", "tests", "gbgpu_amd"):
        shutil.copytree(os.path.join(ROOT, sub), os.path.join(d, sub), ignore=shutil.ignore_patterns("_build", "__pycache__"))
    f = os.path.join(d, "oracle", "fastgb.py")
    s = open(f).read()
    assert s.count(a) == 1, (name, s.count(a))
    open(f, "w").write(s.replace(a, b))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *TESTS], cwd=d, capture_output=True, text=True)
    failed = [l.split(" ")[1] if l.startswith("FAILED") else l for l in r.stdout.splitlines() if l.startswith("FAILED")]
    res.append((name, r.returncode, failed))
    print(f"{name}: rc={r.returncode} first failure: {failed[:1]}", flush=True)
    shutil.rmtree(d)
missed = [n for n, rc, _ in res if rc == 0]
print(f"{len(res) - len(missed)} of {len(res)} mutations caught" + (f"; missed: {missed}" if missed else ""))
sys.exit(1 if missed else 0)
