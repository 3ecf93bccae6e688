"""CPU emulation of the translation products in reduced precision, on the oracle: every
translation matvec (LF M2M/L2L, HF U2U/M2L/D2D) is replaced by an emulated product
while the solves, the kernel applications and the FFT stay FP64.

* FP32: operands rounded to FP32, FP32 accumulation (the mixed mode's MFMA f32 path);
* split-BF16, three terms: each FP32 operand is x = x1 + x2 + x3 with bfloat16 parts
  (8-bit significands, round to nearest even), and the product keeps the six terms
  x1y1 + x1y2 + x2y1 + x1y3 + x3y1 + x2y2, each exact in FP32, accumulated in FP32
  (what six bf16 MFMAs per product would compute);
* split-BF16, two terms (x1y1 + x1y2 + x2y1): the cheaper variant.

* FP32 with three real products per complex one (the mixed-mode GEMM of the e^{ikr}/r
  kernel): k1 = (Ar + Ai) xr, k2 = Ar (xi - xr), k3 = Ai (xr + xi), operand sums and
  products in FP32, Re = k1 - k3, Im = k1 + k2.

The three-term product keeps the FMM as close to the FP64 FMM as FP32 does (about 5e-8
on the 5000-point sphere); the two-term product is about 60 times further away.  The
three-product form stays at the FP32 level for e^{ikr}/r and loses accuracy for e^{ikr}
(larger equivalent densities: the k1 and k3 terms cancel where |Ar| and |Ai| differ
widely), which is why that kernel keeps four products.
"""
import numpy as np
import pytest

from ddfmm_amd import geometry
from oracle import wave3d_ref as R


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest bfloat16 value (round to nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def split(x: np.ndarray, terms: int):
    """Real float32 array -> `terms` bf16 parts whose sum approximates x."""
    parts, rest = [], np.asarray(x, dtype=np.float32)
    for _ in range(terms):
        p = bf16_round(rest)
        parts.append(p)
        rest = (rest - p).astype(np.float32)
    return parts


def _cplx_parts(z: np.ndarray, terms: int):
    z = np.asarray(z, dtype=np.complex64)
    return split(z.real, terms), split(z.imag, terms)


PAIRS = {2: [(0, 0), (0, 1), (1, 0)], 3: [(0, 0), (0, 1), (1, 0), (0, 2), (2, 0), (1, 1)]}


def split_matvec(terms: int):
    def mv(A: np.ndarray, x: np.ndarray) -> np.ndarray:
        Ar, Ai = _cplx_parts(A, terms)
        xr, xi = _cplx_parts(x, terms)
        yr = np.zeros(A.shape[0], dtype=np.float32)
        yi = np.zeros(A.shape[0], dtype=np.float32)
        for i, j in PAIRS[terms]:
            # float32 products of bf16 values are exact; sums accumulate in float32
            yr += (Ar[i] @ xr[j]).astype(np.float32) - (Ai[i] @ xi[j]).astype(np.float32)
            yi += (Ar[i] @ xi[j]).astype(np.float32) + (Ai[i] @ xr[j]).astype(np.float32)
        return (yr + 1j * yi).astype(np.complex128)
    return mv


def fp32_matvec(A: np.ndarray, x: np.ndarray) -> np.ndarray:
    return (np.asarray(A, dtype=np.complex64) @ np.asarray(x, dtype=np.complex64)).astype(np.complex128)


def gauss3_matvec(A: np.ndarray, x: np.ndarray) -> np.ndarray:
    Ar, Ai = np.asarray(A.real, dtype=np.float32), np.asarray(A.imag, dtype=np.float32)
    xr, xi = np.asarray(x.real, dtype=np.float32), np.asarray(x.imag, dtype=np.float32)
    k1 = (Ar + Ai).astype(np.float32) @ xr
    k2 = Ar @ (xi - xr).astype(np.float32)
    k3 = Ai @ (xr + xi).astype(np.float32)
    return ((k1 - k3) + 1j * (k1 + k2)).astype(np.complex128)


def test_gauss3_product_is_the_complex_product():
    rng = np.random.default_rng(1)
    A = (rng.standard_normal((40, 30)) + 1j * rng.standard_normal((40, 30))).astype(np.complex64)
    x = (rng.standard_normal(30) + 1j * rng.standard_normal(30)).astype(np.complex64)
    exact = A.astype(np.complex128) @ x.astype(np.complex128)
    assert np.linalg.norm(gauss3_matvec(A, x) - exact) / np.linalg.norm(exact) < 1e-6


def test_bf16_split_is_exact_decomposition():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(10000).astype(np.float32) * np.float32(1e3)
    p = split(x, 3)
    assert all(np.array_equal(bf16_round(q), q) for q in p)          # each part is a bf16 value
    err = np.abs((p[0].astype(np.float64) + p[1] + p[2]) - x) / np.abs(x)
    assert err.max() < 2.0 ** -22                                      # 3 x 8 bits ~ the FP32 significand
    err2 = np.abs((p[0].astype(np.float64) + p[1]) - x) / np.abs(x)
    assert 2.0 ** -20 < err2.max() < 2.0 ** -15


@pytest.fixture(scope="module")
def problem(ops):
    K, N = 4.0, 5000
    pos = geometry.sphere(N, K, 21)
    den = geometry.densities(N, 21)
    w = R.Wave3dOracle(pos, K, *ops).setup()
    chk = geometry.check_targets(N, 300, 21)
    exact = R.direct_sum(pos[chk], pos, den)
    ref = w.eval(den)
    return w, den, chk, exact, ref


@pytest.mark.parametrize("name,mv,vs_fmm,vs_direct", [
    ("fp32", fp32_matvec, 1e-6, 1e-5),
    ("bf16x3terms", split_matvec(3), 1e-6, 1e-5),
    ("fp32x3products", gauss3_matvec, 1e-6, 1e-5),
])
def test_emulated_translations_meet_the_bar(problem, monkeypatch, name, mv, vs_fmm, vs_direct):
    w, den, chk, exact, ref = problem
    monkeypatch.setattr(R, "translate", mv)
    got = w.eval(den)
    e_fmm = R.relerr(got, ref)
    e_dir = R.relerr(got[chk], exact)
    print(f"{name}: vs FP64 FMM {e_fmm:.2e}, vs direct {e_dir:.2e} (FP64 FMM vs direct {R.relerr(ref[chk], exact):.2e})")
    assert e_fmm < vs_fmm and e_dir < vs_direct


def test_two_term_split_loses_accuracy(problem, monkeypatch):
    w, den, chk, exact, ref = problem
    monkeypatch.setattr(R, "translate", split_matvec(3))
    e3 = R.relerr(w.eval(den), ref)
    monkeypatch.setattr(R, "translate", split_matvec(2))
    e2 = R.relerr(w.eval(den), ref)
    print(f"split-bf16 vs FP64 FMM: 3 terms {e3:.2e}, 2 terms {e2:.2e}")
    assert e2 > 30 * e3


def test_three_products_lose_accuracy_for_exp_ikr(ops_expr, monkeypatch):
    """e^{ikr} (5000-point sphere): three products per complex one move the FMM further
    from the FP64 FMM than four do (1.4 times here; 3 times at config 2 on the GPU)."""
    K, N = 4.0, 5000
    pos = geometry.sphere(N, K, 21)
    den = geometry.densities(N, 21)
    w = R.Wave3dOracle(pos, K, *ops_expr, ktype=1).setup()
    ref = w.eval(den)
    monkeypatch.setattr(R, "translate", fp32_matvec)
    e4 = R.relerr(w.eval(den), ref)
    monkeypatch.setattr(R, "translate", gauss3_matvec)
    e3 = R.relerr(w.eval(den), ref)
    print(f"e^(ikr) vs FP64 FMM: four products {e4:.2e}, three {e3:.2e}")
    assert e3 > 1.2 * e4
