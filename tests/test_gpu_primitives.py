"""Device primitives (prim.hpp) against the reference's own golden blocks and
the oracle, through the C ABI (imago_primitive_batch, imago_boys_batch)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def lm_count(sw):
    return [1, 4, 9, 16][[1, 2, 4, 8].index(sw & 15)], [1, 4, 9, 16][[1, 2, 4, 8].index(sw >> 4)]


@pytest.mark.parametrize("kind", ["ol", "ke", "ep", "np"])
def test_gpu_primitives_vs_reference_golden(gpu_ctx, golden, kind):
    a = golden[kind + "_a"]
    xyz = golden[kind + "_xyz"]
    sw = golden[kind + "_sw"]
    ref = golden[kind + "_sh"]
    print("[gpu] primitive batch", kind, len(sw), flush=True)
    got = gpu_ctx.primitive_batch(kind, a, xyz, sw)
    for t in range(len(sw)):
        na, nb = lm_count(sw[t])
        assert rel_err(got[t][:na, :nb], ref[t][:na, :nb]) < 1e-12, (kind, t, sw[t])
        # entries outside the leading block are written as zero
        mask = np.ones((16, 16), bool)
        mask[:na, :nb] = False
        assert np.all(got[t][mask] == 0.0)


@pytest.mark.parametrize("kind", ["ol", "ke", "ep", "np"])
def test_gpu_primitives_vs_oracle_random(gpu_ctx, oracle_mod, kind):
    rng = np.random.default_rng(11)
    n = 512
    a = np.exp(rng.uniform(np.log(0.1), np.log(5e3), (n, 3)))
    A = rng.normal(size=(n, 3))
    B = A + rng.normal(size=(n, 3)) * 0.8
    C = 0.5 * (A + B) + rng.normal(size=(n, 3)) * 0.5
    sws = np.array([(1 << rng.integers(4)) + (16 << rng.integers(4)) for _ in range(n)])
    print("[gpu] primitive batch random", kind, n, flush=True)
    got = gpu_ctx.primitive_batch(kind, a, np.concatenate([A, B, C], 1), sws)
    worst = 0.0
    for t in range(n):
        ref = oracle_mod.primitive(kind, a[t], A[t], B[t], C[t], sws[t])
        na, nb = lm_count(sws[t])
        if np.max(np.abs(ref[:na, :nb])) < 1e-250:
            continue
        worst = max(worst, rel_err(got[t][:na, :nb], ref[:na, :nb]))
    # NP: the reference Boys closed form loses digits just above T = 0.1
    assert worst < (1e-9 if kind == "np" else 1e-12), worst


def test_gpu_boys_vs_reference_golden(gpu_ctx, golden):
    T = golden["boys_T"]
    F = golden["boys_F"]
    got = gpu_ctx.boys_batch(T)
    series = T <= float(np.float32(0.1))
    # series branch: same float constants, same operations
    assert np.allclose(got[series], F[series], rtol=1e-14, atol=0)
    # closed forms: agreement limited by the cancellation of the formula
    lead_ratio = np.array([[np.sqrt(np.pi) * np.prod(np.arange(1, 2 * m, 2)) / 2 ** (m + 1)
                            * t ** (-m - 0.5) for m in range(7)] for t in T[~series]]) / F[~series]
    err = np.abs(got[~series] - F[~series]) / np.abs(F[~series])
    assert np.all(err <= 1e-13 + 4e-16 * lead_ratio)


def test_primitive_batch_rejects_bad_switch(gpu_ctx):
    from imago_amd.lib import ImagoError
    with pytest.raises(ImagoError) as e:
        gpu_ctx.primitive_batch("ol", np.ones((1, 3)), np.zeros((1, 9)), [3 + 16])
    assert e.value.name == "IMAGO_E_ARG"
