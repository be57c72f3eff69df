"""The register FFTs of the sites kernel (maomi_amd/csrc/fft_reg.hpp), run on the host.

The header is host-and-device code; a small host program built with hipcc runs fft<R> on
seeded inputs and prints the results, which are compared with numpy's DFT.  It also prints the
fused butterflies' twiddle table, whose entries must be the 64th roots of unity to 4.5e-8 with
forward and inverse twiddles exact conjugates, and the energy drift of 240 forward/inverse
round trips, which that pairing keeps small.
"""

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]

HARNESS = r"""
#include "fft_reg.hpp"
#include <cmath>
#include <cstdio>
template <int R, bool INV> void run(FILE* in) {
  float2 v[R];
  for (int i = 0; i < R; ++i) { float a, b; if (fscanf(in, "%f %f", &a, &b) != 2) return; v[i] = make_float2(a, b); }
  fftk::fft<R, INV>(v);
  printf("fft %d %d", R, INV ? 1 : 0);
  for (int i = 0; i < R; ++i) printf(" %.9e %.9e", v[i].x, v[i].y);
  printf("\n");
}
template <int R> void drift(FILE* in) {
  float2 v[R];
  double e0 = 0, e1 = 0;
  for (int i = 0; i < R; ++i) { float a, b; if (fscanf(in, "%f %f", &a, &b) != 2) return; v[i] = make_float2(a, b); e0 += (double)a * a + (double)b * b; }
  for (int r = 0; r < 240; ++r) {
    fftk::fft<R, false>(v);
    fftk::fft<R, true>(v);
    for (int i = 0; i < R; ++i) { v[i].x *= 1.0f / R; v[i].y *= 1.0f / R; }
  }
  for (int i = 0; i < R; ++i) e1 += (double)v[i].x * v[i].x + (double)v[i].y * v[i].y;
  printf("drift %d %.9e\n", R, e1 / e0 - 1.0);
}
int main() {
  for (int e = 0; e < 64; ++e) {
    const fftk::LfTwiddle t = fftk::kLF[e];
    if (t.form == 0) continue;
    const double m = t.m, r = t.r;
    const double c = t.form == 1 ? m : m * r, s = t.form == 1 ? m * r : m;
    printf("tw %d %.17e %.17e\n", e, c, s);
  }
  FILE* in = stdin;
  run<2, false>(in); run<2, true>(in); run<4, false>(in); run<4, true>(in);
  run<8, false>(in); run<8, true>(in); run<16, false>(in); run<16, true>(in);
  run<32, false>(in); run<32, true>(in);
  for (int k = 0; k < 64; ++k) drift<16>(in);
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc is not available")
    d = tmp_path_factory.mktemp("fft_reg")
    src, exe = d / "h.cpp", d / "h"
    src.write_text(HARNESS)
    res = subprocess.run([hipcc, "-O2", "-std=c++17", f"-I{ROOT / 'maomi_amd' / 'csrc'}",
                          f"-I{ROOT / 'include'}", "-Wno-unused-command-line-argument",
                          str(src), "-o", str(exe)], capture_output=True, text=True)
    if res.returncode != 0:
        pytest.fail(f"harness build failed:\n{res.stderr[-2000:]}")
    rng = np.random.default_rng(7)
    sizes = [2, 2, 4, 4, 8, 8, 16, 16, 32, 32]
    inputs = [rng.standard_normal((n, 2)).astype(np.float32) for n in sizes]
    drift_in = [rng.standard_normal((16, 2)).astype(np.float32) for _ in range(64)]
    text = "\n".join(" ".join(f"{a:.9e} {b:.9e}" for a, b in x) for x in inputs + drift_in)
    out = subprocess.run([str(exe)], input=text, capture_output=True, text=True, check=True).stdout
    return inputs, out.splitlines()


def test_register_fft_matches_numpy_dft(harness):
    inputs, lines = harness
    ffts = [ln.split() for ln in lines if ln.startswith("fft ")]
    assert len(ffts) == len(inputs)
    for x, f in zip(inputs, ffts):
        R, inv = int(f[1]), int(f[2])
        got = np.array(f[3:], dtype=np.float64).reshape(R, 2)
        got = got[:, 0] + 1j * got[:, 1]
        z = x[:, 0].astype(np.float64) + 1j * x[:, 1].astype(np.float64)
        ref = np.fft.ifft(z) * R if inv else np.fft.fft(z)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err < 1e-6, (R, inv, err)


def test_twiddle_table_is_conjugate_exact_and_accurate(harness):
    _, lines = harness
    tw = {int(p[1]): complex(float(p[2]), float(p[3]))
          for p in (ln.split() for ln in lines if ln.startswith("tw "))}
    assert sorted(tw) == [e for e in range(64) if e % 16]
    for e, t in tw.items():
        exact = np.exp(2j * np.pi * e / 64)
        assert abs(np.angle(t / exact)) < 5e-8, e
        assert abs(abs(t) ** 2 - 1) < 1e-7, e
        assert tw[64 - e] == t.conjugate(), e  # the inverse uses the forward's conjugate


def test_round_trip_energy_drift_is_small(harness):
    _, lines = harness
    drift = np.array([float(ln.split()[2]) for ln in lines if ln.startswith("drift ")])
    assert len(drift) == 64
    # 240 forward/inverse pairs of 16 points: a bias of the twiddle moduli would compound over
    # the pairs (a modulus off by 1e-7 in one butterfly level gives about 5e-6), rounding averages
    # out over the 64 inputs (measured mean 4.4e-7, largest 5.9e-6)
    assert abs(drift.mean()) < 1e-6 and np.abs(drift).max() < 1e-5
