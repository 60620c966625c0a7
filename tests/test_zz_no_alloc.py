"""quest_run_noalloc (the allocation-free variant of quest_run, the pair sums in
the library's static per-GPU scratch array) against quest_run: same inputs,
same steps, bit-identical states, through the C ABI.  Two calls in a row, the
second on a smaller system (the scratch keeps the first call's contents),
then a batch of uneven systems (its work list has empty items that write a
spare partial row), then a context whose partial sums exceed the scratch
(rejected, nothing run)."""
import numpy as np
import pytest

from conftest import dense_system, short_pulse

pytestmark = pytest.mark.gpu


def _pair(systems_, precision, steps_ref, steps_noalloc, **kw):
    """States after quest_run(sum(steps)) and after quest_run_noalloc(s) for
    each s in steps_noalloc, on two contexts of the same inputs."""
    from quest_amd import Context, systems
    consts = systems.consts()
    pl = short_pulse(area=np.pi / 2, delay=0.006, sigma_t=0.002)
    pulse = [pl] * len(systems_) if isinstance(systems_, list) else pl
    with Context(systems_, consts, precision=precision, pulse=pulse, **kw) as ref:
        ref.run(steps_ref)
        want, dwant = ref.state()
        ref_step = ref.step
    with Context(systems_, consts, precision=precision, pulse=pulse, **kw) as ctx:
        for s in steps_noalloc:
            ctx.run_noalloc(s)
        got, dgot = ctx.state()
        assert ctx.step == ref_step
        assert ctx.info()["graph"] == 0   # no graph was instantiated
    return want, dwant, got, dgot


@pytest.mark.parametrize("precision", [1, 0])
def test_noalloc_matches_run_two_calls_second_smaller(precision):
    for n, side in ((2000, 1.6), (500, 1.0)):
        sysd = dense_system(n, side, seed=40 + n)
        want, dwant, got, dgot = _pair(sysd, precision, 120, [70, 50])
        print(f"precision={precision} n={n}: max|rho01| {np.abs(want[:, 1]).max():.3f}")
        assert np.abs(want[:, 1]).max() > 0.05
        assert np.array_equal(got, want), f"n={n}: states differ"
        assert np.array_equal(dgot, dwant), f"n={n}: derivatives differ"


def test_noalloc_matches_run_on_a_batch():
    sizes = [300, 120, 700, 65, 410]
    batch = [dense_system(n, 0.2 + n / 1000, seed=60 + i) for i, n in enumerate(sizes)]
    want, dwant, got, dgot = _pair(batch, 1, 90, [90])
    assert np.abs(want[:, 1]).max() > 0.05
    assert np.array_equal(got, want)
    assert np.array_equal(dgot, dwant)


def test_noalloc_rejects_partial_sums_beyond_the_scratch():
    from quest_amd import Context, QuestError, systems
    sysd = dense_system(4096, 2.0, seed=5)
    # 2-source chunks: 2049 partial rows x 4096 dots x 16 B = 134 MB > 64 MiB
    with Context(sysd, systems.consts(), precision=1, source_chunk=2,
                 pulse=short_pulse(delay=0.004, sigma_t=0.002)) as ctx:
        with pytest.raises(QuestError, match="static scratch"):
            ctx.run_noalloc(5)
        assert ctx.step == 0
        ctx.run(2)   # the context is untouched and still runs
        assert ctx.step == 2
