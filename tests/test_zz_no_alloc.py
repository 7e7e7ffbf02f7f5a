"""pd_trainer_fit_noalloc (its scratch in a library-owned device array, no
device allocation) against pd_trainer_fit: two trainers with the same seed
and the same samples, fitted call for call, must hold bit-identical losses
and parameters; the second call uses a smaller batch and fewer steps, so it
runs over the scratch the first call left behind.  Requests beyond the
stated maxima fail before any work."""
import sys

import numpy as np
import pytest

from pdhelpers import deformed_pair

pytestmark = pytest.mark.gpu
pds = pytest.importorskip("pdsolver")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_fit_noalloc_equals_fit():
    say("case fit_noalloc vs fit")
    _, ctx = deformed_pair(amp=0.05, s_fixed=0.01)
    kw = dict(dt=1e-4, precond=1, rtol=1e-8)
    ctx.assemble(pds.make_params(**kw))
    ref = pds.Trainer(50000, seed=21)
    tst = pds.Trainer(50000, seed=21)
    na = ref.add(ctx, seed=5, stride=2)
    nb = tst.add(ctx, seed=5, stride=2)
    assert na == nb > 0
    assert np.array_equal(ref.samples(0, na), tst.samples(0, nb))
    for steps, batch, lr, seed in ((40, 1024, 3e-3, 7), (9, 256, 1e-3, 8)):
        say(f"  fit steps {steps} batch {batch}")
        la = ref.fit(steps, batch=batch, lr=lr, seed=seed)
        lb = tst.fit(steps, batch=batch, lr=lr, seed=seed, noalloc=True)
        assert np.all(np.isfinite(lb))
        assert np.array_equal(la, lb), (la[:4], lb[:4])
        assert np.array_equal(ref.params(), tst.params())
    # the fit did change the network
    assert not np.array_equal(tst.params(), pds.Trainer(10, seed=21).params())
    # bounds: rejected with PD_ERR_ARG, state untouched
    before = tst.params()
    with pytest.raises(pds.PDError) as e:
        tst.fit(2, batch=pds.FIT_NOALLOC_MAX_BATCH + 128, noalloc=True)
    assert e.value.code == pds.PD_ERR_ARG
    with pytest.raises(pds.PDError) as e:
        tst.fit(pds.FIT_NOALLOC_MAX_STEPS + 1, batch=128, noalloc=True)
    assert e.value.code == pds.PD_ERR_ARG
    assert np.array_equal(before, tst.params())
    # the largest accepted batch still equals the allocating fit
    la = ref.fit(2, batch=pds.FIT_NOALLOC_MAX_BATCH, lr=1e-3, seed=9)
    lb = tst.fit(2, batch=pds.FIT_NOALLOC_MAX_BATCH, lr=1e-3, seed=9, noalloc=True)
    assert np.array_equal(la, lb) and np.array_equal(ref.params(), tst.params())
    ref.close()
    tst.close()
    ctx.close()
